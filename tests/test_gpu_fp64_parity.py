"""FP64 HIP path vs the CPU oracle, through the C ABI (libnqes.so).

The FP64 device path follows the oracle's canonical operation order, so per-chain
quantities (accept/reject sequence, spins, H_loc, O) must be bit-identical.
Sums over samples use a different (deterministic) reduction order than the
oracle, so moments, S*v and the SR step are compared at 1e-12 relative; the SR
update goes through MINRES-QLP with rtol, so weights are compared at a
tolerance tied to rtol.
"""
import numpy as np
import pytest

from helpers import hs_J, load_J, random_weights, rel_err, xxz_J

pytestmark = pytest.mark.gpu

CASES = [
    # N, M, K, C, m, B
    (8, 16, 2, 6, 4, 3),
    (4, 8, 1, 5, 6, 2),
    (6, 12, 3, 4, 3, 2),
    (4, 4, 4, 3, 3, 2),
    (8, 64, 2, 3, 2, 1),
    (4, 128, 2, 2, 2, 1),
    (40, 80, 2, 2, 2, 1),
]


@pytest.fixture(scope="module")
def nqes():
    import nqes as mod
    return mod


def oracle_samples(oracle, N, M, K, h, J, W, seed, C, m, B, model=0):
    outs = [oracle.run_chain(N, M, K, h, J, W, seed, c, m, B, want_O=True, model=model) for c in range(C)]
    spins = np.concatenate([o["spins"] for o in outs])
    H = np.concatenate([o["H"] for o in outs])
    O = np.concatenate([o["O"] for o in outs])
    acc = np.array([o["accepts"] for o in outs], dtype=np.uint64)
    return spins, H, O, acc


@pytest.mark.parametrize("N,M,K,C,m,B", CASES)
def test_sampler_bitexact(oracle, nqes, N, M, K, C, m, B):
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B}", flush=True)
    h = 0.595
    J = load_J(N)
    W = random_weights(N, M, K, seed=N * 100 + M + K, scale=0.15)
    seed = 1234567
    sp_o, H_o, O_o, acc_o = oracle_samples(oracle, N, M, K, h, J, W, seed, C, m, B)
    with nqes.Solver(N, M, K, h, J) as s:
        s.sample(W, seed, C, m, B, reglam=0.5)
        sp_g, H_g, acc_g = s.samples()
        O_g = s.O_rows(0, C * m)
    assert np.array_equal(acc_g, acc_o), (acc_g, acc_o)
    assert np.array_equal(sp_g, sp_o)
    assert rel_err(H_g, H_o) == 0.0, rel_err(H_g, H_o)
    assert rel_err(O_g, O_o) == 0.0, rel_err(O_g, O_o)


# real weights: O is real, stored as doubles (rows padded to an even length; P = 15
# and P = 2 (3 + 3 + 9) = 30 are odd / not a multiple of four)
REAL_CASES = [
    (8, 16, 2, 6, 4, 3),
    (3, 3, 1, 5, 4, 2),
    (3, 3, 2, 4, 3, 2),
    (6, 12, 3, 4, 3, 2),
]


@pytest.mark.parametrize("N,M,K,C,m,B", REAL_CASES)
def test_sampler_bitexact_real(oracle, nqes, N, M, K, C, m, B):
    """Real weights on the FP64 path: accept sequence, spins, H_loc and the O rows
    (stored as their real parts) bit-identical with the oracle; the near-tie count of
    the accept decisions is reported per chain and is zero."""
    print(f"case real N={N} M={M} K={K} C={C} m={m} B={B}", flush=True)
    h = 0.595
    J = load_J(N)
    W = random_weights(N, M, K, seed=N * 100 + M + K, scale=0.15, complex_=False)
    seed = 1234567
    sp_o, H_o, O_o, acc_o = oracle_samples(oracle, N, M, K, h, J, W, seed, C, m, B)
    assert not np.any(O_o.imag) and not np.any(H_o.imag)
    with nqes.Solver(N, M, K, h, J) as s:
        s.sample(W, seed, C, m, B, reglam=0.5)
        sp_g, H_g, acc_g = s.samples()
        O_g = s.O_rows(0, C * m)
        ties, total = s.ulp_ties()
    assert np.array_equal(acc_g, acc_o), (acc_g, acc_o)
    assert np.array_equal(sp_g, sp_o)
    assert rel_err(H_g, H_o) == 0.0, rel_err(H_g, H_o)
    assert np.array_equal(O_g, O_o)
    assert ties.shape == (C,) and total == 0


# the shapes of the benchmark workloads, at chain counts the oracle finishes in seconds:
# N=32 M=64 K=2 (paper size; one-pass S*v with 4 and 8 waves per row) and N=64 M=128
# K=4 (large; rows too long for the one-pass kernel: two-pass form)
@pytest.mark.parametrize("N,M,K,C,m,B,cplx", [(32, 64, 2, 64, 16, 64, False), (32, 64, 2, 64, 16, 64, True),
                                              (64, 128, 4, 4, 2, 4, True)])
def test_baseline_shapes_bitexact(oracle, nqes, N, M, K, C, m, B, cplx):
    """Spins, accepts, H_loc and O bit-exact with the oracle; moments to 1e-12 and S*v
    (real and complex v) to 1e-12 of the oracle's product over its own O rows."""
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B} cplx={cplx}", flush=True)
    h, reglam = 0.595, 1.0
    J = load_J(N)
    W = random_weights(N, M, K, seed=31 * N + K, scale=0.1 if cplx else 0.05, complex_=cplx)
    seed = 20261203
    ref = oracle.sample_chains(N, M, K, h, J, W, seed, C, m, B, want_O=True)
    O_o = ref["O"]
    S = C * m
    o = O_o.mean(axis=0)
    sq = (np.abs(O_o) ** 2).mean(axis=0) - np.abs(o) ** 2
    with nqes.Solver(N, M, K, h, J) as s:
        s.sample(W, seed, C, m, B, reglam)
        sp_g, H_g, acc_g = s.samples()
        O_g = s.O_rows(0, S)
        mom = s.moments()
        rng = np.random.default_rng(9)
        vr = rng.normal(size=s.P)
        vc = vr + 1j * rng.normal(size=s.P)
        sv = [s.apply_S(v) for v in (vr.astype(np.complex128), vc)]
        _, total = s.ulp_ties()
    assert np.array_equal(acc_g, ref["accepts"])
    assert np.array_equal(sp_g, ref["spins"])
    assert rel_err(H_g, ref["H"]) == 0.0
    assert rel_err(O_g, O_o) == 0.0
    assert rel_err(mom["o"], o) < 1e-12
    assert rel_err(mom["skk"].real, reglam * sq) < 1e-11
    assert total == 0
    for v, out in zip((vr.astype(np.complex128), vc), sv):
        assert rel_err(out, oracle.sv_from_O(O_o, mom["o"], mom["skk"], v)) < 1e-12


@pytest.mark.parametrize("model", [1, 2])
@pytest.mark.parametrize("N,M,K,C,m,B", [(8, 16, 2, 4, 3, 2), (6, 12, 3, 3, 2, 1), (5, 70, 1, 3, 2, 1),
                                        (40, 80, 2, 2, 1, 1)])
def test_sampler_bitexact_pairflip(oracle, nqes, model, N, M, K, C, m, B):
    """hs (model 1) and xxz (model 2) local energies (rbm.cc:1092-1119): spins, H_loc
    and O bit-identical with the oracle (M = 70 puts two hidden units on some lanes)."""
    print(f"case model={model} N={N} M={M} K={K} C={C} m={m} B={B}", flush=True)
    J = hs_J(N) if model == 1 else xxz_J(N)
    W = random_weights(N, M, K, seed=N * 10 + M + K + model, scale=0.15)
    seed = 4242
    sp_o, H_o, O_o, acc_o = oracle_samples(oracle, N, M, K, 0.0, J, W, seed, C, m, B, model=model)
    with nqes.Solver(N, M, K, 0.0, J, model=model) as s:
        s.sample(W, seed, C, m, B, reglam=0.5)
        sp_g, H_g, acc_g = s.samples()
        O_g = s.O_rows(0, C * m)
    assert np.array_equal(acc_g, acc_o)
    assert np.array_equal(sp_g, sp_o)
    assert rel_err(H_g, H_o) == 0.0, rel_err(H_g, H_o)
    assert rel_err(O_g, O_o) == 0.0


@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("N,M,K,C,m,B", CASES[:4])
def test_moments_sv_step(oracle, nqes, N, M, K, C, m, B, cplx):
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B} cplx={cplx}", flush=True)
    h, reglam, lr, rtol = 0.595, 0.7, 0.1, 1e-10
    J = load_J(N)
    W = random_weights(N, M, K, seed=7 + N + M + K, scale=0.15, complex_=cplx)
    seed = 99
    Wo, emat_o, info = oracle.sr_step(N, M, K, h, J, W, seed, C, m, B, lr, reglam, rtol, want_details=True)
    _, _, O_o, _ = oracle_samples(oracle, N, M, K, h, J, W, seed, C, m, B)
    with nqes.Solver(N, M, K, h, J) as s:
        s.sample(W, seed, C, m, B, reglam)
        mom = s.moments()
        assert rel_err(mom["o"], info["o"]) < 1e-12
        assert rel_err(mom["skk"], info["skk"]) < 1e-12
        assert rel_err(mom["f"], info["f"]) < 1e-12
        assert abs(mom["eloc"] - info["eloc"]) <= 1e-12 * abs(info["eloc"])
        assert rel_err(mom["emat"], emat_o) < 1e-12
        rng = np.random.default_rng(5)
        v = rng.normal(size=s.P) + 1j * rng.normal(size=s.P)
        sv_g = s.apply_S(v)
        sv_o = oracle.sv_from_O(O_o, info["o"], info["skk"], v)
        assert rel_err(sv_g, sv_o) < 1e-12
        x, flag, iters = s.solve(rtol)
        assert flag == info["flag"] or flag in (1, 2)
        assert rel_err(x, info["x"]) < 1e-6, (rel_err(x, info["x"]), flag, iters, info["iters"])
        Wg = W.copy()
        emat_g, st = s.monte_carlo(Wg, seed, C, m, B, lr, reglam, rtol, stats=True)
    assert rel_err(emat_g, emat_o) < 1e-12
    assert rel_err(Wg - W, Wo - W) < 1e-6
    assert st["accepts"] == info["accepts"]


@pytest.mark.parametrize("N,M,K,C,m,B", [(8, 16, 2, 16, 8, 4), (6, 12, 3, 8, 6, 2), (4, 8, 1, 10, 5, 2)])
def test_dense_S_matches_matrix_free(nqes, N, M, K, C, m, B):
    """The dense S of the rbm-N* executables (rbm.cc:651-657, 686-696) is the same
    operator as the matrix-free one: S*v equal to 1e-12 on the same samples, and equal
    to skk v + oo v with oo = O^H O / S - conj(o) o^T built in numpy from the O rows;
    one SR step gives the same e_mat and weights to MINRES-QLP rounding."""
    print(f"case N={N} M={M} K={K} C={C} m={m} B={B}", flush=True)
    h, reglam = 0.595, 0.7
    J = load_J(N)
    W = random_weights(N, M, K, seed=3 * N + K, scale=0.15)
    rng = np.random.default_rng(2)
    P = K * (N + M + N * M)
    v = rng.normal(size=P) + 1j * rng.normal(size=P)
    out = {}
    for dense in (False, True):
        with nqes.Solver(N, M, K, h, J) as s:
            if dense:
                s.set_dense_S(True)
            s.sample(W, 55, C, m, B, reglam)
            mom = s.moments()
            O = s.O_rows(0, C * m)
            Wn = W.copy()
            emat = s.monte_carlo(Wn, 55, C, m, B, 0.1, reglam, 1e-11)
            s.sample(W, 55, C, m, B, reglam)
            out[dense] = (s.apply_S(v), emat, Wn)
    o = O.mean(axis=0)
    oo = O.conj().T @ O / (C * m) - np.outer(o.conj(), o)
    ref = mom["skk"] * v + oo @ v
    assert rel_err(out[True][0], ref) < 1e-12
    assert rel_err(out[True][0], out[False][0]) < 1e-12
    assert np.array_equal(out[True][1], out[False][1])
    assert rel_err(out[True][2] - W, out[False][2] - W) < 1e-8
