"""Pins the CPU oracle against independent checks (no GPU needed).

Each oracle function is checked against something that does not restate it:
numpy's independent PCG64, the published golden value, libm via numpy, numpy's
linear algebra, closed forms, exact enumeration of all 2^(K N) multi-spin
configurations against exact diagonalisation, and finite differences.
"""
from fractions import Fraction

import numpy as np
import pytest

from helpers import (dense_hamiltonian, dense_heisenberg, ed_energies, hs_J, load_J, random_weights, rel_err,
                     synthetic_J, xxz_J)

PCG_INC = (6364136223846793005 << 64) | 1442695040888963407
PCG_MULT = (2549297995355413924 << 64) | 4865540595714422341


# ------------------------------------------------------------------ RNG
def test_pcg64_golden(oracle):
    # pcg64(12345)() -- value recorded independently from numpy's PCG64
    assert oracle.PCG64(12345)() == 0xd11b1d37bff50104


@pytest.mark.parametrize("seed", [0, 1, 12345, 3923901102, 2**63 + 5, 2**64 - 1])
def test_pcg64_vs_numpy(oracle, seed):
    bg = np.random.PCG64()
    st = ((seed + PCG_INC) * PCG_MULT + PCG_INC) % (1 << 128)
    bg.state = {"bit_generator": "PCG64", "state": {"state": st, "inc": PCG_INC}, "has_uint32": 0, "uinteger": 0}
    g = oracle.PCG64(seed)
    assert [g() for _ in range(64)] == [int(x) for x in bg.random_raw(64)]


@pytest.mark.parametrize("u", [0, 1, 2**11 - 1, 2**53, 2**53 + 1, 2**63, 2**63 + 1023, 2**63 + 1025,
                               2**64 - 2**11, 2**64 - 2**10 - 1, 2**64 - 2**10, 2**64 - 1])
def test_canonical_closed_form(oracle, u):
    # generate_canonical<double,53> over a 64-bit engine: double(u) (round to nearest
    # even) / 2^64, values that round to 1 are clamped to nextafter(1, 0)
    expected = float(u) / 2.0**64
    if expected >= 1.0:
        expected = np.nextafter(1.0, 0.0)
    assert oracle.canonical(u) == expected
    assert Fraction(oracle.canonical(u)) <= 1


def test_init_spins_bernoulli(oracle):
    # K*N bernoulli(0.5) draws from pcg64(seed + c): replica-major, spin i ascending
    N, K, seed, c = 37, 3, 777, 5
    g = oracle.PCG64(seed + c)
    bits = np.array([[1 if oracle.canonical(g()) < 0.5 else 0 for _ in range(N)] for _ in range(K)])
    assert np.array_equal(oracle.init_spins(N, K, seed, c), oracle.pack_spins(bits))


# ------------------------------------------------------------------ elementary functions
def test_real_functions(oracle):
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.linspace(-700, 700, 4001), rng.normal(size=4000) * 3, [0.0, -0.0, 1e-300, 0.5, 1.0]])
    e = np.array([oracle.exp(x) for x in xs])
    assert np.max(np.abs(e - np.exp(xs)) / np.exp(xs)) < 5e-16
    ys = np.concatenate([np.linspace(-2000, 2000, 4001), rng.normal(size=4000) * 3])
    sc = np.array([oracle.sincos(y) for y in ys])
    assert np.max(np.abs(sc[:, 0] - np.sin(ys))) < 4e-16
    assert np.max(np.abs(sc[:, 1] - np.cos(ys))) < 4e-16
    zs = np.concatenate([np.linspace(-40, 40, 4001), rng.normal(size=4000) * 0.05])
    cs = np.array([oracle.coshsinh(x) for x in zs])
    assert np.max(np.abs(cs[:, 0] - np.cosh(zs)) / np.cosh(zs)) < 5e-16
    nz = np.abs(zs) > 0
    assert np.max(np.abs(cs[nz, 1] - np.sinh(zs[nz])) / np.abs(np.sinh(zs[nz]))) < 5e-16


def test_complex_functions_and_symmetry(oracle):
    rng = np.random.default_rng(1)
    zs = rng.normal(size=3000) * 2 + 1j * rng.normal(size=3000) * 2
    for z in zs:
        for f, ref in ((oracle.cexp, np.exp), (oracle.ccosh, np.cosh), (oracle.csinh, np.sinh),
                       (oracle.ctanh, np.tanh)):
            got, want = f(z), ref(z)
            assert abs(got - want) <= 3e-15 * max(abs(want), 1.0), (f, z, got, want)
        # exact parity: the sampler tabulates cosh(2w), sinh(2w) and flips signs
        assert oracle.ccosh(-z) == oracle.ccosh(z)
        assert oracle.csinh(-z) == -oracle.csinh(z)
    for a, b in zip(zs[:-1], zs[1:]):
        assert abs(oracle.cdiv(a, b) - a / b) <= 3e-16 * abs(a / b) * 4


def test_tree_product(oracle):
    rng = np.random.default_rng(2)
    for M in (1, 5, 16, 64, 65, 100, 130, 512):
        d = np.exp(rng.normal(size=M) * 0.1 + 1j * rng.normal(size=M))
        assert abs(oracle.tree_prod(d) - np.prod(d)) <= 1e-13 * abs(np.prod(d))


# ------------------------------------------------------------------ K x K algebra
@pytest.mark.parametrize("K", [1, 2, 3, 4])
def test_det_inv(oracle, K):
    rng = np.random.default_rng(K)
    for _ in range(50):
        m = rng.normal(size=(K, K)) + 1j * rng.normal(size=(K, K))
        assert abs(oracle.det(m) - np.linalg.det(m)) <= 1e-13 * max(abs(np.linalg.det(m)), 1)
        inv, ok = oracle.inv(m)
        assert ok
        assert rel_err(inv, np.linalg.inv(m)) < 1e-12
    if K > 1:
        m = rng.normal(size=(K, K)) + 0j
        m[1] = m[0]
        inv, ok = oracle.inv(m)
        assert oracle.det(m) == 0 and not ok and not inv.any()


# ------------------------------------------------------------------ amplitudes
def numpy_state(N, M, K, W, bits):
    L = N + M + N * M
    phi = np.zeros((K, K), complex)
    th_all = np.zeros((K, K, M), complex)
    for k in range(K):
        sig = 1 - 2 * bits[k].astype(float)
        for kk in range(K):
            a, b, w = W[kk * L:kk * L + N], W[kk * L + N:kk * L + N + M], W[kk * L + N + M:(kk + 1) * L].reshape(N, M)
            th = b + sig @ w
            phi[k, kk] = np.exp(sig @ a) * np.prod(np.cosh(th))
            th_all[k, kk] = th
    return phi, np.tanh(th_all)


@pytest.mark.parametrize("N,M,K", [(4, 8, 1), (8, 16, 2), (5, 10, 3), (4, 4, 4), (40, 80, 2), (6, 130, 2)])
def test_state_fcn_and_flip(oracle, N, M, K):
    rng = np.random.default_rng(N + M + K)
    W = random_weights(N, M, K, seed=N * M + K, scale=0.2)
    bits = rng.integers(0, 2, size=(K, N))
    sp = oracle.pack_spins(bits)
    phi, tanh = oracle.state_fcn(N, M, K, W, sp)
    phi_np, tanh_np = numpy_state(N, M, K, W, bits)
    assert rel_err(phi, phi_np) < 1e-12
    assert rel_err(tanh, tanh_np) < 1e-12
    # incremental flip == from-scratch amplitude of the flipped configuration
    for k in range(K):
        for i in (0, N // 2, N - 1):
            phi_f, tanh_f = oracle.flip(N, M, K, W, k, i, sp, phi, tanh)
            b2 = bits.copy()
            b2[k, i] ^= 1
            phi2, tanh2 = oracle.state_fcn(N, M, K, W, oracle.pack_spins(b2))
            assert rel_err(phi_f[k], phi2[k]) < 1e-11
            assert rel_err(tanh_f[k], tanh2[k]) < 1e-11
            others = [r for r in range(K) if r != k]
            assert np.array_equal(phi_f[others], phi[others])


def test_ezz(oracle):
    rng = np.random.default_rng(4)
    for N in (2, 8, 33, 64):
        J = load_J(N) if N in (8, 64) else synthetic_J(N)
        for _ in range(20):
            bits = rng.integers(0, 2, size=N)
            sig = 1 - 2 * bits.astype(float)
            assert abs(oracle.ezz(N, J, oracle.pack_spins(bits[None])[0]) - sig @ J @ sig) < 1e-12 * N * N


# ------------------------------------------------------------------ exact enumeration vs ED
def enumerate_all(oracle, N, M, K, h, J, W, model=0):
    """|det Phi|^2-weighted sums over all 2^(K N) multi-spin configurations."""
    tot_w, tot_H = 0.0, np.zeros((K, K), complex)
    Os = []
    for cfg in range(1 << (K * N)):
        bits = np.array([[(cfg >> (k * N + i)) & 1 for i in range(N)] for k in range(K)])
        sp = oracle.pack_spins(bits)
        phi, tanh = oracle.state_fcn(N, M, K, W, sp)
        phi_inv, ok = oracle.inv(phi)
        if not ok:
            continue
        wgt = abs(np.linalg.det(phi)) ** 2
        H = oracle.hloc(N, M, K, h, J, W, sp, phi, phi_inv, tanh, model=model)
        tot_w += wgt
        tot_H += wgt * H
        Os.append((wgt, H, oracle.deriv(N, M, K, sp, phi, phi_inv, tanh)))
    return tot_H / tot_w, Os, tot_w


def rbm_vectors(N, M, K, W):
    """columns psi_kk(s) over the 2^N single-replica basis."""
    L = N + M + N * M
    s = np.arange(1 << N)
    sig = 1 - 2 * ((s[:, None] >> np.arange(N)[None, :]) & 1).astype(float)
    out = np.zeros((1 << N, K), complex)
    for kk in range(K):
        a, b, w = W[kk * L:kk * L + N], W[kk * L + N:kk * L + N + M], W[kk * L + N + M:(kk + 1) * L].reshape(N, M)
        out[:, kk] = np.exp(sig @ a) * np.prod(np.cosh(b[None, :] + sig @ w), axis=1)
    return out


@pytest.mark.parametrize("N,M,K", [(6, 6, 1), (4, 4, 2), (3, 3, 3)])
def test_hloc_expectation_equals_rayleigh_ritz(oracle, N, M, K):
    # E_{|det Phi|^2}[Phi^-1 (H Phi)] = G^-1 A^H H A with A = [psi_0 .. psi_{K-1}],
    # G = A^H A: the estimator the sampler averages is the Rayleigh-Ritz matrix of H
    # on span{psi_k}.  H is the dense sum_{i!=j} J_ij Z_i Z_j - h sum_i X_i.
    h = 0.595
    J = synthetic_J(N, 1.0) * 0.9
    W = random_weights(N, M, K, seed=11 * N + K, scale=0.3)
    Hmean, _, _ = enumerate_all(oracle, N, M, K, h, J, W)
    A = rbm_vectors(N, M, K, W)
    Hd = dense_hamiltonian(N, h, J)
    G = A.conj().T @ A
    R = np.linalg.solve(G, A.conj().T @ Hd @ A)
    assert rel_err(Hmean, R) < 1e-10, (Hmean, R)
    # a transposed estimator would not pass
    if K > 1:
        assert rel_err(Hmean, R.T) > 1e-3


@pytest.mark.parametrize("model,delta", [(1, 1.0), (2, -1.0)])
@pytest.mark.parametrize("N,M,K", [(6, 6, 1), (4, 4, 2), (3, 3, 3)])
def test_hs_hloc_expectation_equals_rayleigh_ritz(oracle, model, delta, N, M, K):
    # the same identity for the pair-flip local energy of rbm.cc:1092-1119 with the
    # dense sum_{i>j} J_ij (XX + YY + delta ZZ); random dense couplings exercise every
    # pair, the model couplings (hs: all pairs, xxz: ring bonds) the shipped ones
    rng = np.random.default_rng(5 * N + K + model)
    W = random_weights(N, M, K, seed=13 * N + K + model, scale=0.3)
    A = rbm_vectors(N, M, K, W)
    G = A.conj().T @ A
    for J in (np.tril(rng.normal(size=(N, N)), -1), hs_J(N) if model == 1 else xxz_J(N)):
        Hmean, _, _ = enumerate_all(oracle, N, M, K, 0.3, J, W, model=model)
        R = np.linalg.solve(G, A.conj().T @ dense_heisenberg(N, J, delta) @ A)
        assert rel_err(Hmean, R) < 1e-10, (Hmean, R)
        if K > 1:
            assert rel_err(Hmean, R.T) > 1e-3


def test_hs_xxz_couplings_and_spectra():
    # hs_model / xxz_model (rbm.cc:1066-1089) and textbook spectra of the chains they
    # define: Haldane-Shastry ground state E0 = -(pi^2/24)(N + 5/N) in units of
    # sum J (S.S) with S = sigma/2 (here H = 4 sum J S.S), and the XXZ ring at
    # delta = -1 is -sum (XX + YY - ZZ): by the sublattice rotation equal to the
    # Heisenberg antiferromagnet sum (XX + YY + ZZ), whose N = 4 ground state is -8
    for N in (4, 6, 8):
        E0 = np.linalg.eigvalsh(dense_heisenberg(N, hs_J(N), 1.0))[0]
        assert abs(E0 - 4 * (-(np.pi ** 2 / 24) * (N + 5.0 / N))) < 1e-9, (N, E0)
    assert abs(np.linalg.eigvalsh(dense_heisenberg(4, xxz_J(4), -1.0))[0] + 8.0) < 1e-12
    J = xxz_J(6)
    assert J[1, 0] == -1.0 and J[5, 0] == -1.0 and J[5, 4] == -1.0 and np.count_nonzero(J) == 6


def test_ed_matches_published_values():
    # E/N of the two lowest states of the N=8 trapped-ion chain (J from the
    # reference's Ising2d_N8_alpha1.0.mat); values computed independently by ED
    J = load_J(8)
    assert np.allclose(ed_energies(8, 1.0, J, 2), [-1.534182, -1.502611], atol=2e-6)
    assert np.allclose(ed_energies(8, 0.595, J, 2), [-1.342304, -1.341163], atol=2e-6)
    assert np.allclose(ed_energies(8, 0.5, J, 2), [-1.315832, -1.315530], atol=2e-6)


def test_deriv_finite_differences(oracle):
    # O_p = d log det Phi / d W_p (holomorphic), central differences along Re W_p
    N, M, K = 4, 8, 2
    W = random_weights(N, M, K, seed=5, scale=0.2)
    rng = np.random.default_rng(6)
    bits = rng.integers(0, 2, size=(K, N))
    sp = oracle.pack_spins(bits)
    phi, tanh = oracle.state_fcn(N, M, K, W, sp)
    phi_inv, _ = oracle.inv(phi)
    O = oracle.deriv(N, M, K, sp, phi, phi_inv, tanh)
    eps = 1e-6
    fd = np.zeros_like(O)
    for p in range(len(W)):
        Wp, Wm = W.copy(), W.copy()
        Wp[p] += eps
        Wm[p] -= eps
        dp = np.linalg.det(oracle.state_fcn(N, M, K, Wp, sp)[0])
        dm = np.linalg.det(oracle.state_fcn(N, M, K, Wm, sp)[0])
        fd[p] = np.log(dp / dm) / (2 * eps)
    assert rel_err(O, fd) < 1e-7


# ------------------------------------------------------------------ S*v and MINRES-QLP
def test_sv_matches_dense_S(oracle):
    rng = np.random.default_rng(8)
    S, P = 300, 40
    O = rng.normal(size=(S, P)) + 1j * rng.normal(size=(S, P))
    o = O.mean(axis=0)
    skk = 0.3 * ((np.abs(O) ** 2).mean(axis=0) - np.abs(o) ** 2) + 0j
    v = rng.normal(size=P) + 1j * rng.normal(size=P)
    Sd = (O.conj().T @ O) / S - np.outer(o.conj(), o) + np.diag(skk)
    assert rel_err(oracle.sv_from_O(O, o, skk, v), Sd @ v) < 1e-13


def test_minresqlp_vs_direct(oracle):
    rng = np.random.default_rng(9)
    n = 40
    X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = X.conj().T @ X + 0.5 * np.eye(n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    x, flag, it = oracle.minresqlp_dense(A, b, 1e-12)
    assert flag in (1, 3)
    assert rel_err(x, np.linalg.solve(A, b)) < 1e-8
    # singular, consistent system: MINRES-QLP returns the minimum-length solution
    U = rng.normal(size=(n, 10)) + 1j * rng.normal(size=(n, 10))
    As = U @ U.conj().T
    bs = As @ (rng.normal(size=n) + 0j)
    xs, flag, _ = oracle.minresqlp_dense(As, bs, 1e-12)
    assert rel_err(xs, np.linalg.pinv(As) @ bs) < 1e-6
    # b = 0 -> x = 0, flag 0
    x0, flag0, it0 = oracle.minresqlp_dense(A, np.zeros(n, complex), 1e-12)
    assert flag0 == 0 and it0 == 0 and not x0.any()


def test_sr_solution_solves_dense_system(oracle):
    # the SR step's x solves (S_dense + diag(skk)) x = f built from its own samples
    N, M, K, h, C, m, B = 4, 8, 2, 0.8, 8, 20, 5
    J = synthetic_J(N)
    W = random_weights(N, M, K, seed=12, scale=0.1)
    seed = 31
    _, _, info = oracle.sr_step(N, M, K, h, J, W, seed, C, m, B, 0.1, 2.0, 1e-12, want_details=True)
    O = np.concatenate([oracle.run_chain(N, M, K, h, J, W, seed, c, m, B, want_O=True)["O"] for c in range(C)])
    S = C * m
    o = O.mean(axis=0)
    assert rel_err(o, info["o"]) < 1e-12
    A = (O.T * 1.0) @ O.conj() / S
    A = A.T - np.outer(o.conj(), o) + np.diag(info["skk"])
    assert rel_err(A @ info["x"], info["f"]) < 1e-8


def test_sr_loop_reaches_ed(oracle):
    # a short training run of N=4 (schedule of the small configuration: lr 0.1/0.99,
    # reglam 40/0.9 cut 1, rtol 1e-9) reaches the exact two lowest levels.  Trained
    # much longer, this tiny system's weights keep growing until Phi overflows -- the
    # reference's --throw-fperror/exit-127 regime, covered by the non-finite tests.
    N, M, K, h = 4, 8, 2, 0.595
    J = synthetic_J(N)
    W = random_weights(N, M, K, seed=3, scale=0.01, complex_=False)
    lr, reglam = 0.1, 40.0
    hist = []
    for it in range(90):
        W, emat, info = oracle.sr_step(N, M, K, h, J, W, 1000 + it * 64, 64, 100, 32, max(lr, 0.01),
                                       max(reglam, 1.0), 1e-9)
        lr *= 0.99
        reglam *= 0.9
        hist.append(np.sort(np.linalg.eigvals(emat / N).real))
    ev = np.mean(hist[-5:], axis=0)
    assert np.allclose(ev, ed_energies(N, h, J, K), atol=3e-3), (ev, ed_energies(N, h, J, K))
    # real weights stay real through the complex algorithm
    assert not np.any(W.imag)


# ------------------------------------------------------------------ correlation operators
def test_correlation_operators_rayleigh_ritz(oracle):
    # E_{|det Phi|^2}[Phi^+ fcn(idx)] = G^-1 A^H Op A for Op = sigma^x_i, sigma^z_i,
    # sigma^x_i sigma^x_j, sigma^z_i sigma^z_j; fcn built as rbm.cc:774-829 does
    N, M, K = 4, 4, 2
    W = random_weights(N, M, K, seed=77, scale=0.3)
    A = rbm_vectors(N, M, K, W)
    G = A.conj().T @ A
    dim = 1 << N
    s = np.arange(dim)
    sig = 1 - 2 * ((s[:, None] >> np.arange(N)[None, :]) & 1)
    ops = {}
    for i in range(N):
        X = np.zeros((dim, dim))
        X[s, s ^ (1 << i)] = 1
        ops[("x", i)] = X
        ops[("z", i)] = np.diag(sig[:, i]).astype(float)
    acc = {key: np.zeros((K, K), complex) for key in ops}
    acc_xx = np.zeros((N, N, K, K), complex)
    tot = 0.0
    for cfg in range(1 << (K * N)):
        bits = np.array([[(cfg >> (k * N + i)) & 1 for i in range(N)] for k in range(K)])
        phi, _ = oracle.state_fcn(N, M, K, W, oracle.pack_spins(bits))
        phi_inv, ok = oracle.inv(phi)
        if not ok:
            continue
        wgt = abs(np.linalg.det(phi)) ** 2
        tot += wgt
        for i in range(N):
            b2 = bits.copy()
            b2[:, i] ^= 1
            phx, _ = oracle.state_fcn(N, M, K, W, oracle.pack_spins(b2))
            acc[("x", i)] += wgt * phi_inv @ phx
            acc[("z", i)] += wgt * phi_inv @ (np.where(bits[:, i:i + 1] == 1, -1, 1) * phi)
            for j in range(N):
                b3 = bits.copy()
                b3[:, i] ^= 1
                b3[:, j] ^= 1
                phxx, _ = oracle.state_fcn(N, M, K, W, oracle.pack_spins(b3))
                acc_xx[i, j] += wgt * phi_inv @ phxx
    for key, Op in ops.items():
        R = np.linalg.solve(G, A.conj().T @ Op @ A)
        assert rel_err(acc[key] / tot, R) < 1e-10, key
    for i in range(N):
        for j in range(N):
            R = np.linalg.solve(G, A.conj().T @ (ops[("x", i)] @ ops[("x", j)]) @ A)
            assert rel_err(acc_xx[i, j] / tot, R) < 1e-10
    # the oracle's sampler-side estimator uses exactly these per-configuration
    # matrices: on a chain, its xx_ii is the identity and xx, zz are symmetric
    x, z, xx, zz = oracle.correlations(N, M, K, 0.5, synthetic_J(N), W, 5, 3, 6, 2)
    assert np.abs(xx[np.arange(N), np.arange(N)] - np.eye(K)).max() < 1e-12
    assert np.array_equal(xx, xx.transpose(1, 0, 2, 3)) and np.array_equal(zz, zz.transpose(1, 0, 2, 3))


def test_sample_chains_equals_run_chain(oracle):
    """orc_sample_chains_model (chains in parallel) returns exactly the per-chain
    orc_run_chain_model results, in chain order."""
    from helpers import load_J, random_weights
    N, M, K, C, m, B = 8, 16, 2, 7, 4, 3
    J = load_J(N)
    W = random_weights(N, M, K, seed=12, scale=0.15)
    a = oracle.sample_chains(N, M, K, 0.595, J, W, 31, C, m, B, want_O=True)
    per = [oracle.run_chain(N, M, K, 0.595, J, W, 31, c, m, B, want_O=True) for c in range(C)]
    assert np.array_equal(a["spins"], np.concatenate([r["spins"] for r in per]))
    assert np.array_equal(a["H"], np.concatenate([r["H"] for r in per]))
    assert np.array_equal(a["O"], np.concatenate([r["O"] for r in per]))
    assert list(a["accepts"]) == [r["accepts"] for r in per]
