"""Exhaustive evaluation (nqes_enumerate) and the --dump-states/ham/deriv files of the
rbm-N* executables (rbm.cc:474-526, 1149-1172).

* Phi of every multi-spin configuration equals the oracle's state_fcn bit for bit
  (same canonical FP64 order); the |det Phi|^2-weighted mean of Phi^+ res over all
  configurations is the Rayleigh-Ritz matrix G^-1 A^H H A of the dense Hamiltonian
  (trapped-ion and xxz), an independent check of the enumerated local energies.
* The executable's dump files: variable s, complex, dims [2^(K N), rows, cols]
  column-major; H_loc and O equal the oracle's (Cramer inverse) on nonsingular Phi to
  1e-10, and numpy's pseudo-inverse form on the singular configurations (two replicas
  on the same spins), as the reference's pseudo-inverse gives.
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio

from helpers import dense_hamiltonian, dense_heisenberg, random_weights, rel_err, xxz_J

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "nqes-trapped-ion_amd", "bin", "rbm-matfree")


def rbm_vectors(N, M, K, W):
    L = N + M + N * M
    s = np.arange(1 << N)
    sig = 1 - 2 * ((s[:, None] >> np.arange(N)[None, :]) & 1).astype(float)
    out = np.zeros((1 << N, K), complex)
    for kk in range(K):
        a, b, w = W[kk * L:kk * L + N], W[kk * L + N:kk * L + N + M], W[kk * L + N + M:(kk + 1) * L].reshape(N, M)
        out[:, kk] = np.exp(sig @ a) * np.prod(np.cosh(b[None, :] + sig @ w), axis=1)
    return out


def cfg_spins(oracle, cfg, N, K):
    bits = np.array([[(cfg >> (k * N + i)) & 1 for i in range(N)] for k in range(K)])
    return oracle.pack_spins(bits)


@pytest.mark.parametrize("model,N,M,K", [(0, 5, 5, 2), (0, 6, 12, 1), (2, 4, 8, 2), (1, 3, 6, 3)])
def test_enumeration_bitexact_and_rayleigh_ritz(oracle, model, N, M, K):
    import nqes
    print(f"case model={model} N={N} M={M} K={K}", flush=True)
    h = 0.595
    if model == 0:
        J = 0.3 * (np.ones((N, N)) - np.eye(N))  # symmetric, E_zz = 2 sum_{i>j}
    elif model == 1:
        J = 0.7 * np.tril(np.ones((N, N)), -1)   # lower triangle, every pair
    else:
        J = xxz_J(N)
    W = random_weights(N, M, K, seed=17 + N + K + model, scale=0.2)
    with nqes.Solver(N, M, K, h, J, model=model) as s:
        phi, res, tanh = s.enumerate(W)
    ncfg = 1 << (N * K)
    for cfg in range(0, ncfg, max(1, ncfg // 40)):
        sp = cfg_spins(oracle, cfg, N, K)
        phi_o, tanh_o = oracle.state_fcn(N, M, K, W, sp)
        assert np.array_equal(phi[cfg], phi_o), cfg
        assert np.array_equal(tanh[cfg], tanh_o.reshape(K, K, M)), cfg
    wgt = np.abs(np.linalg.det(phi)) ** 2
    H = np.einsum("nij,njk->nik", np.linalg.pinv(phi), res)
    Hmean = np.einsum("n,nij->ij", wgt, H) / wgt.sum()
    A = rbm_vectors(N, M, K, W)
    Hd = dense_hamiltonian(N, h, J) if model == 0 else dense_heisenberg(N, J, 1.0 if model == 1 else -1.0)
    R = np.linalg.solve(A.conj().T @ A, A.conj().T @ Hd @ A)
    assert rel_err(Hmean, R) < 1e-9, (Hmean, R)


def test_cli_dump_files(oracle, tmp_path):
    N, M, K = 4, 8, 2
    W = random_weights(N, M, K, seed=23, scale=0.2)
    L = N + M + N * M
    Wk = W.reshape(K, L)
    sio.savemat(str(tmp_path / "w0.mat"), {"a": Wk[:, :N], "b": Wk[:, N:N + M], "w": Wk[:, N + M:].reshape(K, N, M)})
    J = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            if i != j:
                J[i, j] = abs(i - j) ** -1.0
    J /= J.max()
    sio.savemat(str(tmp_path / "J.mat"), {"J": J})
    link = tmp_path / "rbm-N4-M8-K2-opt"
    os.symlink(BIN, link)
    cmd = [str(link), "--model", "trappedion", "-h", "0.5", "-c", "4", "-m", "2", "-n", "0", "--seed", "3",
           "--jmat", str(tmp_path / "J.mat"), "--load-rbm", str(tmp_path / "w0.mat"),
           "--correlation-data", str(tmp_path / "c.mat"), "--dump-states", str(tmp_path / "st"),
           "--dump-ham", str(tmp_path / "ham"), "--dump-deriv", str(tmp_path / "der")]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    st = sio.loadmat(str(tmp_path / "stN4M8K2.mat"))["s"]
    ham = sio.loadmat(str(tmp_path / "hamN4M8K2.mat"))["s"]
    der = sio.loadmat(str(tmp_path / "derN4M8K2.mat"))["s"]
    ncfg, P = 1 << (N * K), K * L
    assert st.shape == (ncfg, K, K) and ham.shape == (ncfg, K, K) and der.shape == (ncfg, P)
    import nqes
    with nqes.Solver(N, M, K, 0.5, J) as s:
        _, res, _ = s.enumerate(W)
    nsing = 0
    for cfg in range(ncfg):
        sp = cfg_spins(oracle, cfg, N, K)
        phi_o, tanh_o = oracle.state_fcn(N, M, K, W, sp)
        assert np.array_equal(st[cfg], phi_o)
        inv, ok = oracle.inv(phi_o)
        sv = np.linalg.svd(phi_o, compute_uv=False)
        if sv[-1] > 1e-10 * sv[0]:
            H_o = oracle.hloc(N, M, K, 0.5, J, W, sp, phi_o, inv, tanh_o)
            assert rel_err(ham[cfg], H_o) < 1e-10, cfg
            O_o = oracle.deriv(N, M, K, sp, phi_o, inv, tanh_o)
            assert rel_err(der[cfg], O_o) < 1e-9, cfg
        else:
            nsing += 1
            # pseudo-inverse form: H = Phi^+ res, O = det of a rank-deficient matrix
            assert rel_err(ham[cfg], np.linalg.pinv(phi_o) @ res[cfg]) < 1e-10, cfg
            assert np.all(np.abs(der[cfg]) < 1e-8)
    assert nsing >= 1 << N  # every configuration with equal replicas
