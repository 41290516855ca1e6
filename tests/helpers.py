"""Shared test helpers: J fixtures, synthetic weights, exact diagonalisation."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_J(N):
    """J of the trapped-ion chain from the reference's Jmat fixtures (copied as .npy)."""
    path = os.path.join(GOLDEN, f"J_N{N}.npy")
    if os.path.exists(path):
        return np.load(path)
    return synthetic_J(N)


def synthetic_J(N, alpha=1.0):
    """J_ij = |i-j|^-alpha normalised to max 1, zero diagonal."""
    i = np.arange(N)
    d = np.abs(i[:, None] - i[None, :]).astype(float)
    J = np.zeros((N, N))
    nz = d > 0
    J[nz] = d[nz] ** (-alpha)
    return J / J.max() if N > 1 else J


def random_weights(N, M, K, seed, scale=0.1, complex_=True):
    rng = np.random.default_rng(seed)
    P = K * (N + M + N * M)
    W = rng.normal(size=P) * scale
    if complex_:
        W = W + 1j * rng.normal(size=P) * scale
    return W.astype(np.complex128)


def dense_hamiltonian(N, h, J):
    """H = sum_{i!=j} J_ij Z_i Z_j - h sum_i X_i on 2^N states (bit i set = sigma_i = -1)."""
    dim = 1 << N
    s = np.arange(dim)
    sig = 1 - 2 * ((s[:, None] >> np.arange(N)[None, :]) & 1)
    diag = np.einsum("si,ij,sj->s", sig, J, sig)
    H = np.diag(diag).astype(float)
    for i in range(N):
        H[s, s ^ (1 << i)] -= h
    return H


def hs_J(N):
    """Haldane-Shastry couplings (lower triangle): J_ij = 1/d^2, d = (N/pi) sin(pi (i-j)/N)."""
    J = np.zeros((N, N))
    for i in range(1, N):
        for j in range(i):
            d = np.sin(np.pi * (i - j) / N) * N / np.pi
            J[i, j] = 1.0 / (d * d)
    return J


def xxz_J(N):
    """XXZ chain couplings (lower triangle): -1 on the bonds (i, i-1) and (N-1, 0)."""
    J = np.zeros((N, N))
    for i in range(1, N):
        J[i, i - 1] = -1.0
    if N > 2:
        J[N - 1, 0] = -1.0
    return J


def dense_heisenberg(N, J, delta):
    """H = sum_{i>j} J_ij (X_i X_j + Y_i Y_j + delta Z_i Z_j) on 2^N states, J's lower
    triangle (bit i set = sigma_i = -1); XX + YY maps |..01..> to 2 |..10..>."""
    dim = 1 << N
    H = np.zeros((dim, dim))
    for s in range(dim):
        for i in range(1, N):
            for j in range(i):
                p = J[i, j]
                if p == 0.0:
                    continue
                bi, bj = (s >> i) & 1, (s >> j) & 1
                if bi == bj:
                    H[s, s] += p * delta
                else:
                    H[s, s] -= p * delta
                    H[s ^ (1 << i) ^ (1 << j), s] += 2.0 * p
    return H


def ed_energies(N, h, J, K):
    return np.linalg.eigvalsh(dense_hamiltonian(N, h, J))[:K] / N


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)
