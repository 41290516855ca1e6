"""Host-side parts of bench.py: the exact-diagonalisation energies its gates use, and
the one-GPU guard (no device needed)."""
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_ed_sparse_equals_dense():
    """The Lanczos branch (N > 10) returns the same K lowest levels as dense
    diagonalisation, including the lowest level of the odd sector of the global spin
    flip, which a flip-even start vector never reaches."""
    N, h, K = 11, 0.595, 3
    J = bench.power_law_J(N, 1.0)
    dense = bench.ed_energies(N, h, J, K, dense=True)
    sparse = bench.ed_energies(N, h, J, K, dense=False)
    np.testing.assert_allclose(sparse, dense, rtol=0, atol=1e-12)
    # the two lowest levels lie in different flip sectors (ordered phase at this h)
    dim = 1 << N
    s = np.arange(dim)
    sig = 1.0 - 2.0 * ((s[:, None] >> np.arange(N)[None, :]) & 1)
    H = np.diag(np.einsum("si,ij,sj->s", sig, J, sig))
    for i in range(N):
        H[s, s ^ (1 << i)] -= h
    w, V = np.linalg.eigh(H)
    flip = dim - 1 - s  # all spins flipped
    parity = [np.sign(V[flip, k] @ V[:, k]) for k in range(2)]
    assert parity[0] != parity[1]
    np.testing.assert_allclose(w[:K] / N, dense, rtol=0, atol=1e-13)


def test_ed_n8_values():
    """E/N at N=8, h=0.595 with the ion chain's J (the values the N=8 gate uses)."""
    ed = bench.ed_energies(8, 0.595, bench.load_J(8), 2)
    np.testing.assert_allclose(ed, [-1.342304, -1.341163], atol=5e-7)


def test_more_than_one_gpu_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout
