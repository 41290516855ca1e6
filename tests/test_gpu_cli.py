"""End-to-end checks of the drop-in executable.

* rbm-matfree (installed under the reference's name rbm-matfree-N8-M16-K2-opt)
  prints the seed line and one line of ascending E/N values per SR step in the
  reference's format, identical to the Python mirror's SR loop on the same weights;
  writes --correlation-data (x, z: N x K; xx, zz: N x N x K) and --dump-rbm files
  that scipy reads; exits 127 with the reference's message on a floating-point
  error under --throw-fperror.
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.io as sio

from helpers import hs_J, load_J, random_weights, xxz_J

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "nqes-trapped-ion_amd", "bin", "rbm-matfree")
GOLDEN = os.path.join(REPO, "tests", "golden")


def weights_to_mat(path, W, N, M, K):
    L = N + M + N * M
    Wk = W.reshape(K, L)
    sio.savemat(path, {"a": Wk[:, :N], "b": Wk[:, N:N + M], "w": Wk[:, N + M:].reshape(K, N, M)})


def mat_to_weights(path, N, M, K):
    d = sio.loadmat(path)
    return np.concatenate([d["a"], d["b"], d["w"].reshape(K, N * M)], axis=1).reshape(-1)


@pytest.fixture()
def exe(tmp_path):
    link = tmp_path / "rbm-matfree-N8-M16-K2-opt"
    os.symlink(BIN, link)
    return str(link)


def run_cli(exe, tmp_path, extra, check=True):
    J = load_J(8)
    sio.savemat(str(tmp_path / "J8.mat"), {"J": J})
    cmd = [exe, "--model", "trappedion", "-h", "0.595", "-c", "16", "-m", "20", "--seed", "77",
           "--reglam-init", "40", "--reglam-cutoff", "1", "--minres-rtol", "1e-9",
           "--correlation-data", str(tmp_path / "corr.mat"), "--jmat", str(tmp_path / "J8.mat")] + extra
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    if check:
        assert r.returncode == 0, r.stderr
    return r


def test_cli_matches_python_mirror_and_writes_files(exe, tmp_path):
    import nqes
    N, M, K = 8, 16, 2
    W = random_weights(N, M, K, seed=5, scale=0.1)
    weights_to_mat(str(tmp_path / "w0.mat"), W, N, M, K)
    r = run_cli(exe, tmp_path, ["-n", "3", "--load-rbm", str(tmp_path / "w0.mat"), "--dump-rbm",
                                str(tmp_path / "out")])
    lines = r.stdout.splitlines()
    assert lines[0] == "Random seed is 77"
    printed = [[float(x) for x in ln.split()] for ln in lines[1:4]]
    with nqes.Solver(N, M, K, 0.595, load_J(N)) as s:
        Wp = W.copy()
        hist, conv, _ = nqes.run(s, Wp, 77, 16, 20, 64, 3, reglam0=40, reglam_cutoff=1, rtol=1e-9)
    for a, b in zip(printed, hist):
        assert np.allclose(a, b, rtol=1e-5, atol=0), (a, b)
    # Eigen's default formatting: 6 significant digits, right aligned to the widest
    for ln, ev in zip(lines[1:4], hist):
        strs = ["%g" % v for v in ev]
        w = max(len(x) for x in strs)
        assert ln == " ".join(x.rjust(w) for x in strs)
    Wd = mat_to_weights(str(tmp_path / "outN8M16K2.mat"), N, M, K)
    assert np.allclose(Wd, Wp, rtol=0, atol=1e-15)
    c = sio.loadmat(str(tmp_path / "corr.mat"))
    assert c["x"].shape == (N, K) and c["z"].shape == (N, K)
    assert c["xx"].shape == (N, N, K) and c["zz"].shape == (N, N, K)
    # <sigma^x_i sigma^x_i> = <sigma^z_i sigma^z_i> = 1 for every state
    assert np.allclose(c["xx"][np.arange(N), np.arange(N)], 1.0, atol=1e-10)
    assert np.allclose(c["zz"][np.arange(N), np.arange(N)], 1.0, atol=1e-10)


@pytest.mark.parametrize("model,mid", [("xxz", 2), ("hs", 1)])
def test_cli_pairflip_models_match_mirror(exe, tmp_path, model, mid):
    """--model xxz / hs: couplings of rbm.cc:1066-1089 built by the executable and the
    pair-flip local energy; the energy lines equal the Python mirror's loop over a
    Solver with the test's own hs_J / xxz_J."""
    import nqes
    N, M, K = 8, 16, 2
    W = random_weights(N, M, K, seed=8, scale=0.1)
    weights_to_mat(str(tmp_path / "w0.mat"), W, N, M, K)
    cmd = [exe, "--model", model, "-h", "0", "-c", "16", "-m", "20", "--seed", "77", "-n", "3",
           "--reglam-init", "40", "--reglam-cutoff", "1", "--minres-rtol", "1e-9",
           "--correlation-data", str(tmp_path / "corr.mat"), "--load-rbm", str(tmp_path / "w0.mat")]
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    printed = [[float(x) for x in ln.split()] for ln in r.stdout.splitlines()[1:4]]
    J = hs_J(N) if model == "hs" else xxz_J(N)
    with nqes.Solver(N, M, K, 0.0, J, model=mid) as s:
        hist, _, _ = nqes.run(s, W.copy(), 77, 16, 20, 64, 3, reglam0=40, reglam_cutoff=1, rtol=1e-9)
    assert len(printed) == 3
    for a, b in zip(printed, hist):
        assert np.allclose(a, b, rtol=1e-5, atol=0), (a, b)


def test_cli_dense_name_selects_dense_S(tmp_path):
    """Installed as rbm-N8-M16-K2-opt (the reference's dense-S build) the executable
    solves with the dense S: same samples, so the first energy line is identical to
    the matrix-free executable's and later lines agree to MINRES-QLP rounding."""
    N, M, K = 8, 16, 2
    W = random_weights(N, M, K, seed=9, scale=0.1)
    weights_to_mat(str(tmp_path / "w0.mat"), W, N, M, K)
    outs = {}
    for name in ("rbm-N8-M16-K2-opt", "rbm-matfree-N8-M16-K2-opt"):
        link = tmp_path / name
        os.symlink(BIN, link)
        r = run_cli(str(link), tmp_path, ["-n", "3", "--load-rbm", str(tmp_path / "w0.mat"), "--stats"])
        outs[name] = [[float(x) for x in ln.split()] for ln in r.stdout.splitlines()[1:4]]
    a, b = outs["rbm-N8-M16-K2-opt"], outs["rbm-matfree-N8-M16-K2-opt"]
    assert a[0] == b[0]
    assert np.allclose(a, b, rtol=1e-6, atol=0)


def test_cli_gauge_and_single_rbm(exe, tmp_path):
    N, M, K = 8, 16, 2
    W = random_weights(N, M, K, seed=6, scale=0.1)
    L = N + M + N * M
    weights_to_mat(str(tmp_path / "w0.mat"), W, N, M, K)
    run_cli(exe, tmp_path, ["-n", "0", "--load-rbm", str(tmp_path / "w0.mat"), "--apply-gauge", "--dump-rbm",
                            str(tmp_path / "g")])
    Wg = mat_to_weights(str(tmp_path / "gN8M16K2.mat"), N, M, K)
    # loaded with a_i -= i pi/2 on odd i, dumped with another -= i pi/2
    d = (Wg - W).reshape(K, L)
    assert np.allclose(d[:, 1:N:2], -1j * np.pi, atol=1e-14) and np.allclose(d[:, 0:N:2], 0)
    assert np.allclose(d[:, N:], 0)
    sio.savemat(str(tmp_path / "s.mat"), {"a": W[:N][None], "b": W[N:N + M][None],
                                          "w": W[N + M:L].reshape(1, N, M)})
    run_cli(exe, tmp_path, ["-n", "0", "--load-single-rbm", str(tmp_path / "s.mat"), "--dump-rbm",
                            str(tmp_path / "s")])
    Ws = mat_to_weights(str(tmp_path / "sN8M16K2.mat"), N, M, K).reshape(K, L)
    assert np.array_equal(Ws[0], W[:L]) and np.array_equal(Ws[1, :N], -W[:N])
    assert np.array_equal(Ws[1, N:N + M], W[N:N + M]) and np.array_equal(Ws[1, N + M:], -W[N + M:L])


def test_cli_fperror_exit_127(exe, tmp_path):
    N, M, K = 8, 16, 2
    W = np.full(K * (N + M + N * M), 500.0, dtype=np.complex128)  # cosh(1000) overflows
    weights_to_mat(str(tmp_path / "big.mat"), W, N, M, K)
    r = run_cli(exe, tmp_path, ["-n", "2", "--load-rbm", str(tmp_path / "big.mat"), "--throw-fperror"], check=False)
    assert r.returncode == 127
    assert "Floating point error occured. Exiting." in r.stderr
