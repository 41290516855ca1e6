"""Mixed-precision HIP path (FP32 amplitudes/O, FP64 sums) through the C ABI.

Tolerances:
  * the mixed path draws the same pcg64 streams as the FP64 path; an accept
    decision differs only when the FP32 amplitude ratio lands within ~1e-6 of the
    uniform draw, so at these sizes nearly all chains keep identical spins;
  * on chains with identical spins, H_loc and O agree with the FP64 path to 1e-4
    relative (FP32 per-hidden-unit arithmetic, FP64 products and sums);
  * S*v on the mixed path's own FP32 O equals the dense (1/S) O^H O - o* o^T +
    diag(skk) product built in float64 from the same O to 1e-10 (FP64 accumulation)
    with the streamed pass over O (NQES_SV=stored);
  * the factored S*v (default: FP64 MFMA passes over the spins and the FP32 factor
    rows c t) differs from it only by the FP32 rounding of O's sum over replicas,
    so it agrees to 1e-6 relative (1e-12 for K = 1, where that sum has one term);
  * e_mat and the SR update agree with the FP64 path to 1e-4 relative.
"""
import numpy as np
import pytest

from helpers import load_J, random_weights, rel_err, hs_J, xxz_J

pytestmark = pytest.mark.gpu

CASES = [
    # N, M, K, C, m, B, complex weights
    (8, 16, 2, 64, 8, 4, False),
    (8, 16, 2, 64, 8, 4, True),
    (6, 12, 1, 32, 8, 4, True),
    (6, 12, 3, 32, 6, 4, False),
    (4, 8, 4, 16, 6, 4, True),
    (16, 64, 2, 32, 4, 2, False),
    (32, 64, 2, 32, 4, 2, False),
    (16, 160, 2, 16, 3, 2, True),
    (40, 80, 2, 16, 3, 2, False),
    (80, 80, 2, 16, 3, 2, False),  # three spin words per replica (streamed S*v above 64 spins)
    (72, 72, 1, 16, 3, 2, True),
    (64, 128, 4, 16, 3, 2, False),  # the large workload's shape (S reduced: FP64 O is 26 MB here)
]


@pytest.fixture(scope="module")
def nqes():
    import nqes as mod
    return mod


def run_both(nqes, N, M, K, cplx, C, m, B, seed=777, reglam=0.5):
    J = load_J(N)
    W = random_weights(N, M, K, seed=N + M + K, scale=0.1, complex_=cplx)
    out = {}
    for prec in (nqes.PREC_FP64, nqes.PREC_MIXED):
        with nqes.Solver(N, M, K, 0.595, J, precision=prec) as s:
            s.sample(W, seed, C, m, B, reglam)
            sp, H, acc = s.samples()
            out[prec] = dict(spins=sp, H=H, acc=acc, O=s.O_rows(0, C * m), mom=s.moments())
    return W, J, out


@pytest.mark.parametrize("N,M,K,C,m,B,cplx", CASES)
def test_mixed_matches_fp64(nqes, N, M, K, C, m, B, cplx):
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B} complex={cplx}", flush=True)
    W, J, out = run_both(nqes, N, M, K, cplx, C, m, B)
    a, b = out[nqes.PREC_FP64], out[nqes.PREC_MIXED]
    S = C * m
    same = np.all(a["spins"].reshape(S, -1) == b["spins"].reshape(S, -1), axis=1)
    assert same.mean() >= 0.9, same.mean()
    assert rel_err(b["H"][same], a["H"][same]) < 1e-4
    assert rel_err(b["O"][same], a["O"][same]) < 1e-4
    if same.all():
        assert rel_err(b["mom"]["emat"], a["mom"]["emat"]) < 1e-4
        assert rel_err(b["mom"]["f"], a["mom"]["f"]) < 1e-3
    if not cplx:
        assert not np.any(b["O"].imag)


@pytest.mark.parametrize("N,M,K,C,m,B,cplx", [CASES[0], CASES[1], CASES[3], CASES[6], CASES[8]])
def test_tables_in_lds_or_global_same_bits(nqes, monkeypatch, N, M, K, C, m, B, cplx):
    """The sweeps and the measurement read their padded cosh/sinh, w, b and exp(+-2a)
    tables from LDS when they fit and from global memory otherwise (NQES_TABLDS=0
    forces the latter); the arithmetic is the same, so spins, H_loc, O and the moments
    agree bit for bit."""
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B} complex={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=N + 3 * M + K, scale=0.1, complex_=cplx)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("NQES_TABLDS", mode)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 4242, C, m, B, 0.5)
            sp, H, acc = s.samples()
            out[mode] = (sp, H, acc, s.O_rows(0, C * m), s.moments())
    a, b = out["1"], out["0"]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[3], b[3])
    for key in ("o", "skk", "f", "emat"):
        assert np.array_equal(a[4][key], b[4][key]), key


@pytest.fixture()
def sv_stored(monkeypatch):
    monkeypatch.setenv("NQES_SV", "stored")


@pytest.mark.parametrize("N,M,K,C,m,B,cplx", [CASES[0], CASES[1], CASES[3], CASES[6], CASES[7]])
def test_mixed_sv_exact_on_own_O(nqes, sv_stored, N, M, K, C, m, B, cplx):
    print(f"case N={N} M={M} K={K} complex={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=3 + N, scale=0.1, complex_=cplx)
    reglam = 0.3
    with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
        s.sample(W, 99, C, m, B, reglam)
        O = s.O_rows(0, C * m)
        mom = s.moments()
        S = C * m
        o = O.mean(axis=0)
        assert rel_err(mom["o"], o) < 1e-12
        skk = reglam * ((np.abs(O) ** 2).mean(axis=0) - np.abs(o) ** 2)
        assert rel_err(mom["skk"].real, skk) < 1e-10
        rng = np.random.default_rng(1)
        for v in (rng.normal(size=s.P) + 0j, rng.normal(size=s.P) + 1j * rng.normal(size=s.P)):
            Sv = O.conj().T @ (O @ v) / S - o.conj() * (o @ v) + mom["skk"] * v
            assert rel_err(s.apply_S(v), Sv) < 1e-10


@pytest.mark.parametrize("N,M,K,C,m,B,cplx", [
    (8, 16, 2, 64, 8, 4, False), (8, 16, 2, 64, 8, 4, True), (6, 12, 1, 32, 8, 4, True),
    (6, 12, 3, 32, 6, 4, False), (8, 16, 4, 16, 6, 4, True), (32, 64, 2, 64, 4, 2, False),
    (40, 80, 2, 16, 3, 2, True), (64, 64, 3, 8, 3, 2, False), (17, 34, 4, 9, 5, 2, True)])
def test_factored_sv_matches_stored(nqes, monkeypatch, N, M, K, C, m, B, cplx):
    """Samples whose Phi is singular (two replicas on the same spins) give unbounded
    Cramer weights c, so O is ill-defined there in both forms (the reference's
    pseudo-inverse would give O = det(Phi^+ derPhi) = 0).  With 2^N spin states per
    replica that needs tiny N; K = 4 cases use N >= 8."""
    print(f"case N={N} M={M} K={K} C={C} m={m} complex={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=11 + N, scale=0.1, complex_=cplx)
    rng = np.random.default_rng(5)
    vs = [rng.normal(size=K * (N + M + N * M)) + 0j,
          rng.normal(size=K * (N + M + N * M)) + 1j * rng.normal(size=K * (N + M + N * M))]
    out, mom, Orows = {}, {}, {}
    for mode in ("fp64mfma", "stored"):
        monkeypatch.setenv("NQES_SV", mode)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 4321, C, m, B, 0.7)
            out[mode] = [s.apply_S(v) for v in vs]
            mom[mode] = s.moments()
            Orows[mode] = s.O_rows(0, C * m)
    tol = 1e-12 if K == 1 else 1e-6
    out["factored"], mom["factored"], Orows["factored"] = out["fp64mfma"], mom["fp64mfma"], Orows["fp64mfma"]
    for a, b in zip(out["factored"], out["stored"]):
        assert rel_err(a, b) < tol, rel_err(a, b)
    # O rows rebuilt from the factors equal the stored rows to FP32 rounding (the
    # kernels may contract c * tanh(theta) into an FMA differently on the two paths)
    assert rel_err(Orows["factored"], Orows["stored"]) < 1e-6
    # moments from the factors (FP64 on the exact factors) vs from the FP32 O rows
    O = Orows["stored"]
    o = O.mean(axis=0)
    for key in ("o", "skk", "f"):
        assert rel_err(mom["factored"][key], mom["stored"][key]) < tol, (key, rel_err(mom["factored"][key], mom["stored"][key]))
    assert rel_err(mom["factored"]["o"], o) < tol
    skk = 0.7 * ((np.abs(O) ** 2).mean(axis=0) - np.abs(o) ** 2)
    assert rel_err(mom["factored"]["skk"].real, skk) < max(tol, 1e-10)
    assert mom["factored"]["emat"] is not None and np.array_equal(mom["factored"]["emat"], mom["stored"]["emat"])


@pytest.mark.parametrize("N,M,K,C,m", [(8, 16, 1, 37, 5), (8, 16, 2, 64, 8), (17, 34, 3, 29, 7), (32, 64, 2, 96, 9),
                                        (40, 80, 2, 33, 5), (64, 128, 4, 21, 3), (24, 24, 4, 50, 4)])
def test_q8_sv_matches_fp64_mfma(nqes, monkeypatch, N, M, K, C, m):
    """The int8 form of the factored S*v (real weights) against the FP64-MFMA form on
    the same samples: equal up to the quantisation of the factors to 2^-24 of each
    row's largest entry (the size of the FP32 rounding of the factors themselves;
    measured 1.3e-7 to 2.5e-7 relative, bound 5e-7), linear and symmetric to 1e-11 (v
    and the row weights are quantised to 2^-47 and 2^-41 of their scale, the weights
    per block of rows).  Row counts S K that are not multiples of 64 exercise the
    padding; N = 40 and 64 the 64-spin MFMA, N = 64 M = 128 K = 4 several column
    groups in both passes."""
    print(f"case N={N} M={M} K={K} C={C} m={m}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=7 + N, scale=0.1, complex_=False)
    P = K * (N + M + N * M)
    rng = np.random.default_rng(9)
    v1, v2 = rng.normal(size=P) + 0j, rng.normal(size=P) * 1e-3 + 0j
    out = {}
    for mode in ("fp64mfma", "q8"):
        if mode == "q8":
            monkeypatch.delenv("NQES_SV", raising=False)
        else:
            monkeypatch.setenv("NQES_SV", mode)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 77, C, m, 3, 0.5)
            out[mode] = [s.apply_S(v) for v in (v1, v2, v1 + 2.5 * v2)]
    a1, a2, a12 = out["q8"]
    b1, b2, _ = out["fp64mfma"]
    assert rel_err(a1, b1) < 5e-7, rel_err(a1, b1)
    assert rel_err(a2, b2) < 5e-7, rel_err(a2, b2)
    assert rel_err(a12, a1 + 2.5 * a2) < 1e-11, rel_err(a12, a1 + 2.5 * a2)
    sym = abs(np.vdot(v2, a1) - np.vdot(v1, a2)) / (np.linalg.norm(v1) * np.linalg.norm(a2))
    assert sym < 1e-11, sym


@pytest.mark.parametrize("N,M,K,C,m", [(8, 16, 1, 37, 5), (8, 16, 2, 64, 8), (17, 34, 3, 29, 7), (32, 64, 2, 96, 9),
                                        (40, 80, 2, 33, 5), (64, 128, 4, 21, 3), (24, 24, 4, 50, 4)])
def test_q8c_sv_matches_fp64_mfma(nqes, monkeypatch, N, M, K, C, m):
    """Complex weights on the int8 form: each complex factor row is quantised with one
    scale into three digit sets, X_R, X_I and X_A = X_R + X_I (exact integers), and S*v
    runs Gauss's three real products each way (forward: A on Re v, R on Im v - Re v, I on
    Re v + Im v; backward: A with Re w + Im w, R with Im w, I with Re w), the three sets'
    small kernels (v digits, row weights, ordered reductions) in one launch each.
    Against the FP64-MFMA form on the same samples: equal up to the factor quantisation
    (2^-24 of each row's largest |Re| + |Im|; bound 5e-7), linear to 1e-11 and Hermitian:
    <u, S v> = conj(<v, S u>) to 1e-11."""
    print(f"case complex N={N} M={M} K={K} C={C} m={m}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=9 + N, scale=0.1, complex_=True)
    P = K * (N + M + N * M)
    rng = np.random.default_rng(19)
    v1 = rng.normal(size=P) + 1j * rng.normal(size=P)
    v2 = (rng.normal(size=P) + 1j * rng.normal(size=P)) * 1e-3
    out = {}
    for mode in ("fp64mfma", "q8"):
        if mode == "q8":
            monkeypatch.delenv("NQES_SV", raising=False)
        else:
            monkeypatch.setenv("NQES_SV", mode)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 79, C, m, 3, 0.5)
            out[mode] = [s.apply_S(v) for v in (v1, v2, v1 + (2.5 - 0.5j) * v2, v1.real + 0j)]
    a1, a2, a12, ar = out["q8"]
    b1, b2, _, br = out["fp64mfma"]
    for a, b in ((a1, b1), (a2, b2), (ar, br)):
        assert rel_err(a, b) < 5e-7, rel_err(a, b)
    assert rel_err(a12, a1 + (2.5 - 0.5j) * a2) < 1e-11, rel_err(a12, a1 + (2.5 - 0.5j) * a2)
    herm = abs(np.vdot(v2, a1) - np.conj(np.vdot(v1, a2))) / (np.linalg.norm(v1) * np.linalg.norm(a2))
    assert herm < 1e-11, herm


@pytest.mark.parametrize("N,M,K,C,m", [(8, 16, 1, 37, 5), (8, 16, 2, 64, 8), (12, 24, 2, 45, 7), (16, 64, 2, 33, 9),
                                        (16, 16, 4, 40, 5), (32, 64, 2, 96, 9), (24, 24, 4, 50, 4), (32, 32, 2, 70, 11)])
def test_q8_one_pass_matches_two_pass(nqes, monkeypatch, N, M, K, C, m):
    """The one-pass int8 S*v (real weights, N <= 32, K in {1, 2, 4}, K M <= 128: the
    factor digits read once, the forward products, the row weights and the backward
    products of each 64-row chunk in one kernel) against the two-pass form
    (NQES_Q8_FUSED=0) on the same samples: the same factor digits and the same
    quantisation of v; the one-pass form quantises the row weights with a running scale
    (at most 2^2 above the largest weight so far) instead of one per block of rows, and
    adds the a block (the K columns c) in FP64 instead of on int8 MFMA, so the two agree
    to ~1e-13 relative (bound 1e-11).  The moments (the same backward products with the
    weights 1 and E) likewise; row counts S K that are not multiples of 128 exercise the
    padding, and a v with one block 1e-3 of the rest the running scale."""
    print(f"case N={N} M={M} K={K} C={C} m={m}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=21 + N + K, scale=0.1, complex_=False)
    P = K * (N + M + N * M)
    rng = np.random.default_rng(23)
    v1 = rng.normal(size=P) + 0j
    v2 = v1.copy()
    v2[: P // 3] *= 1e-3
    out = {}
    monkeypatch.delenv("NQES_SV", raising=False)
    for mode in ("1", "0"):
        monkeypatch.setenv("NQES_Q8_FUSED", mode)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 81, C, m, 3, 0.5)
            out[mode] = ([s.apply_S(v) for v in (v1, v2, v1 + 2.5 * v2)], s.moments())
    (a1, a2, a12), ma = out["1"]
    (b1, b2, _), mb = out["0"]
    assert rel_err(a1, b1) < 1e-11, rel_err(a1, b1)
    assert rel_err(a2, b2) < 1e-11, rel_err(a2, b2)
    assert rel_err(a12, a1 + 2.5 * a2) < 1e-11, rel_err(a12, a1 + 2.5 * a2)
    sym = abs(np.vdot(v2, a1) - np.vdot(v1, a2)) / (np.linalg.norm(v1) * np.linalg.norm(a2))
    assert sym < 1e-11, sym
    for key in ("o", "skk", "f", "emat"):
        assert rel_err(ma[key], mb[key]) < 1e-11, (key, rel_err(ma[key], mb[key]))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("N,M,K,C,m", [(8, 16, 1, 37, 5), (17, 34, 3, 29, 7), (32, 64, 2, 96, 9), (24, 24, 4, 50, 4)])
def test_q8_forward_passes_agree(nqes, monkeypatch, N, M, K, C, m, cplx):
    """For N <= 32 the int8 forward pass runs as one GEMM over the columns (v quantised
    with one scale for all of it, the bias as digit slots of an extra spin tile); the
    tiled form (per-column scales, NQES_Q8_FWD=tiles) computes the same products of the
    same factor digits.  Both quantise v to 2^-47 of a scale, so S*v agrees to ~1e-13
    relative (bound 1e-11); v with blocks of very different size (1e-3) included."""
    print(f"case N={N} M={M} K={K} C={C} m={m} complex={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=11 + N, scale=0.1, complex_=cplx)
    P = K * (N + M + N * M)
    rng = np.random.default_rng(13)
    v1 = rng.normal(size=P) + (1j * rng.normal(size=P) if cplx else 0j)
    v2 = v1.copy()
    v2[: P // 3] *= 1e-3
    out = {}
    monkeypatch.delenv("NQES_SV", raising=False)
    for mode in ("gemm", "tiles"):
        if mode == "tiles":
            monkeypatch.setenv("NQES_Q8_FWD", "tiles")
        else:
            monkeypatch.delenv("NQES_Q8_FWD", raising=False)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 78, C, m, 3, 0.5)
            out[mode] = [s.apply_S(v) for v in (v1, v2)]
    for a, b in zip(out["gemm"], out["tiles"]):
        assert rel_err(a, b) < 1e-11, rel_err(a, b)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("N,M,K,C,m", [(8, 16, 1, 37, 5), (12, 24, 2, 45, 7), (32, 64, 2, 96, 9), (24, 24, 4, 50, 4),
                                        (16, 64, 4, 20, 3)])
def test_q8_c_columns_apart_agree(nqes, monkeypatch, N, M, K, C, m, cplx):
    """Two-pass GEMM form with K M a multiple of 16 (K in {1, 2, 4}): the K c columns of
    each factor row stay quantised integers outside the digit planes, their forward
    products and the a block of the backward pass in FP64 with the unquantised v and
    row weights; NQES_Q8_CSPLIT=0 keeps them as an extra digit column tile.  The same
    factor integers either way, v and the weights quantised to 2^-47 and 2^-41 of their
    scale in the digit form, so S*v agrees to ~1e-13 relative (bound 1e-11), and each
    form stays linear and symmetric (Hermitian) to 1e-11.  K M = 256 (N = 16, M = 64,
    K = 4) takes the GEMM forward pass only with the c columns apart (four 64-column
    groups); the other way it runs the tiled forward pass."""
    print(f"case N={N} M={M} K={K} C={C} m={m} complex={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=31 + N + K, scale=0.1, complex_=cplx)
    P = K * (N + M + N * M)
    rng = np.random.default_rng(29)
    v1 = rng.normal(size=P) + (1j * rng.normal(size=P) if cplx else 0j)
    v2 = v1.copy()
    v2[: P // 3] *= 1e-3
    out = {}
    monkeypatch.delenv("NQES_SV", raising=False)
    monkeypatch.setenv("NQES_Q8_FUSED", "0")
    for mode in ("1", "0"):
        monkeypatch.setenv("NQES_Q8_CSPLIT", mode)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 83, C, m, 3, 0.5)
            out[mode] = [s.apply_S(v) for v in (v1, v2, v1 + 2.5 * v2)]
    for a, b in zip(out["1"], out["0"]):
        assert rel_err(a, b) < 1e-11, rel_err(a, b)
    for mode in ("1", "0"):
        a1, a2, a12 = out[mode]
        assert rel_err(a12, a1 + 2.5 * a2) < 1e-11, (mode, rel_err(a12, a1 + 2.5 * a2))
        sym = abs(np.vdot(v2, a1) - np.conj(np.vdot(v1, a2))) / (np.linalg.norm(v1) * np.linalg.norm(a2))
        assert sym < 1e-11, (mode, sym)


@pytest.mark.parametrize("cplx", [False, True])
def test_q8_graph_replays_equal_eager(nqes, monkeypatch, cplx):
    """The int8 S*v runs as a captured HIP graph (one per sample set; for complex
    weights the two forward and two backward passes over the split rows): every replay,
    with eager work and synchronisations between replays and a second sample set
    (graph rebuilt), equals the eagerly launched sequence (NQES_GRAPH=0) bit for bit."""
    N, M, K, C, m = 16, 32, 2, 64, 8
    J = load_J(16)
    W = random_weights(N, M, K, seed=41, scale=0.1, complex_=cplx)
    P = K * (N + M + N * M)
    rng = np.random.default_rng(3)
    vs = [rng.normal(size=P) + 0j for _ in range(4)] + [rng.normal(size=P) + 1j * rng.normal(size=P)]
    out = {}
    for graph in ("1", "0"):
        monkeypatch.setenv("NQES_GRAPH", graph)
        monkeypatch.delenv("NQES_SV", raising=False)
        res = []
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            for seed in (5, 6):
                s.sample(W, seed, C, m, 2, 0.5)
                for v in vs:
                    res.append(s.apply_S(v))
                    s.moments()
                    s.samples()
            Wn = W.copy()
            res.append(s.monte_carlo(Wn, 7, C, m, 2, 0.1, 1.0, 1e-9))
            res.append(Wn)
        out[graph] = res
    for a, b in zip(out["1"], out["0"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("model", [1, 2])
@pytest.mark.parametrize("cplx", [False, True])
def test_mixed_pairflip_matches_fp64(nqes, model, cplx):
    """hs / xxz local energies on the mixed path (FP32 angles and pair-flip ratios,
    FP64 sums) against the FP64 path on the same chains: H_loc samples to 1e-4, one SR
    step's energies to 1e-4."""
    N, M, K, C, m, B = 10, 20, 2, 32, 8, 4
    J = hs_J(N) if model == 1 else xxz_J(N)
    W = random_weights(N, M, K, seed=31 + model, scale=0.1, complex_=cplx)
    res = {}
    for prec in (nqes.PREC_FP64, nqes.PREC_MIXED):
        with nqes.Solver(N, M, K, 0.0, J, precision=prec, model=model) as s:
            s.sample(W, 17, C, m, B, 1.0)
            sp, H, _ = s.samples()
            Wn = W.copy()
            emat = s.monte_carlo(Wn, 17, C, m, B, 0.05, 2.0, 1e-9)
            res[prec] = (sp, H, emat)
    (sa, Ha, ea), (sb, Hb, eb) = res[nqes.PREC_FP64], res[nqes.PREC_MIXED]
    same = np.all(sa.reshape(C * m, -1) == sb.reshape(C * m, -1), axis=1)
    assert same.mean() > 0.9
    assert rel_err(Hb[same], Ha[same]) < 1e-4
    assert rel_err(eb, ea) < 1e-4


@pytest.mark.parametrize("cplx", [False, True])
def test_mixed_sr_step_matches_fp64(nqes, cplx):
    N, M, K, C, m, B = 8, 16, 2, 64, 16, 8
    J = load_J(N)
    W = random_weights(N, M, K, seed=21, scale=0.1, complex_=cplx)
    res = {}
    for prec in (nqes.PREC_FP64, nqes.PREC_MIXED):
        with nqes.Solver(N, M, K, 0.595, J, precision=prec) as s:
            Wn = W.copy()
            emat, st = s.monte_carlo(Wn, 5, C, m, B, 0.1, 2.0, 1e-9, stats=True)
            res[prec] = (Wn, emat, st)
    (Wa, ea, sa), (Wb, eb, sb) = res[nqes.PREC_FP64], res[nqes.PREC_MIXED]
    assert rel_err(eb, ea) < 1e-4
    assert rel_err(Wb - W, Wa - W) < 1e-3
    assert abs(int(sa["accepts"]) - int(sb["accepts"])) <= max(3, sa["accepts"] // 1000)
    if not cplx:
        assert not np.any(Wb.imag)


@pytest.mark.parametrize("N,M,K,C,cplx", [(8, 16, 2, 64, True), (32, 64, 2, 128, False), (40, 80, 2, 64, True),
                                          (64, 128, 4, 32, True), (17, 34, 3, 50, False)])
def test_theta_gemm_forms_agree(nqes, monkeypatch, N, M, K, C, cplx):
    """The chains' starting angles Theta0 = Sigma W + b on BF16 MFMA (W split exactly into
    three bf16 terms, FP32 accumulation; the default) and on FP32 MFMA (NQES_THETA=fp32)
    are the same exact products summed in FP32 in another order, so they differ by a few
    ulps; the first sweep's amplitude ratios then differ by ~1e-7 and decide a proposal
    differently only when a draw falls that close: the sampled spins and accept counts
    agree on at least 99 % of the chains and H_loc on the chains that agree to 1e-5."""
    print(f"case theta N={N} M={M} K={K} C={C} complex={cplx}", flush=True)
    J = load_J(N)
    # weights of scale 0.1: at 0.3 the FP32 angle arithmetic itself (either theta form)
    # already decides many proposals differently from FP64 at N = 64, M = 128
    W = random_weights(N, M, K, seed=3 + N, scale=0.1, complex_=cplx)
    out = {}
    for form in ("bf16", "fp32"):
        if form == "fp32":
            monkeypatch.setenv("NQES_THETA", "fp32")
        else:
            monkeypatch.delenv("NQES_THETA", raising=False)
        with nqes.Solver(N, M, K, 0.595, J, precision=nqes.PREC_MIXED) as s:
            s.sample(W, 2024, C, 4, 1, 0.5)
            out[form] = s.samples()
    (sa, Ha, aa), (sb, Hb, ab) = out["bf16"], out["fp32"]
    same_chain = np.all((sa == sb).reshape(C, -1), axis=1) & (aa == ab)
    assert same_chain.mean() >= 0.99, same_chain.mean()
    rows = np.repeat(same_chain, 4)
    assert rel_err(Ha[rows], Hb[rows]) < 1e-5
