"""The library's host MINRES-QLP (csrc/minresqlp.hpp, as the SR step runs it: operator
in launch and finish halves, update passes fused with the reductions that follow them,
buffers rotated instead of copied) against the oracle's restatement of minresqlp.h on
dense Hermitian systems: the same flag, iteration count and solution, bit for bit.  The
systems cover positive definite, singular (the QLP steps), norm-limited (the MINRES
step of the first iteration: A scaled by 1e-10) and iteration-capped solves, complex
and real.  The oracle's own solution is pinned against numpy in test_oracle.py."""
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def solver(tmp_path_factory):
    exe = tmp_path_factory.mktemp("mq") / "minres_dense"
    csrc = os.path.join(REPO, "nqes-trapped-ion_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I", csrc, os.path.join(REPO, "tests", "cpp", "minres_dense.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def run_lib(solver, tmp_path, A, b, rtol, maxit):
    n = len(b)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([n], dtype=np.int64).tofile(f)
        np.array([rtol], dtype=np.float64).tofile(f)
        np.array([maxit], dtype=np.uint64).tofile(f)
        np.ascontiguousarray(A, dtype=np.complex128).tofile(f)
        np.ascontiguousarray(b, dtype=np.complex128).tofile(f)
    subprocess.run([solver, str(src), str(dst)], check=True)
    raw = np.fromfile(dst, dtype=np.uint8)
    flag = int(raw[:8].view(np.int64)[0])
    iters = int(raw[8:16].view(np.uint64)[0])
    x = raw[16:].view(np.complex128)
    return x, flag, iters


def systems():
    rng = np.random.default_rng(21)
    out = []
    for t in range(12):
        n = 24 + 9 * t
        cplx = t % 2 == 1
        rank = n // 2 if t % 4 == 3 else n
        B = rng.normal(size=(n, rank)) + (1j * rng.normal(size=(n, rank)) if cplx else 0)
        shift = 0.0 if t % 3 == 0 else 1e-3 * (1 + t % 5)
        A = B @ B.conj().T + shift * np.eye(n)
        if t % 5 == 4:
            A = A * 1e-10  # solution norm past 1e7: the norm limit trips in the first iteration
        b = rng.normal(size=n) + (1j * rng.normal(size=n) if cplx else 0)
        out.append((t, A, b))
    return out


@pytest.mark.parametrize("t,A,b", systems(), ids=lambda p: str(p) if isinstance(p, int) else "")
def test_library_minresqlp_equals_oracle(solver, oracle, tmp_path, t, A, b):
    n = len(b)
    for rtol in (1e-6, 1e-12):
        for maxit in (3, n, 4 * n):
            x, flag, iters = run_lib(solver, tmp_path, A, b, rtol, maxit)
            xo, flago, itero = oracle.minresqlp_dense(A, b, rtol, maxit)
            assert (flag, iters) == (flago, itero), (t, rtol, maxit)
            assert np.array_equal(x.view(np.uint64), xo.view(np.uint64)), (t, rtol, maxit)
