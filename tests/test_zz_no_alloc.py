"""nqes_apply_S_rows_noalloc, the S*v entry point that allocates nothing, against
nqes_apply_S on the same samples (FP64 path, 1e-12 relative: the two run the same
kernels, so they agree up to the final combine, done on the host by one and on the
device by the other).

The caller holds O, o, skk, v and the output in device memory and passes a stream;
this test allocates them through the HIP runtime libnqes.so itself links (ctypes on
the process's own libamdhip64, so one runtime serves both), and the library's
partial rows live in its static device array.  Two calls in a row per kind of O
(complex, real), the second on a smaller problem, each checked; rows too long for the
one-pass kernel, NULL operands and a short row stride are refused.
"""
import ctypes

import numpy as np
import pytest

from helpers import load_J, random_weights, rel_err

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """The HIP runtime libnqes.so is linked against, found in this process's maps."""

    def __init__(self, nqes):
        nqes.load_library()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    path = line.split()[-1]
                    break
        assert path, "libamdhip64 not loaded with libnqes.so"
        self.rt = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [vp]
        self.rt.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        self.rt.hipStreamSynchronize.argtypes = [vp]
        self.rt.hipStreamDestroy.argtypes = [vp]
        assert self.rt.hipSetDevice(0) == 0
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def env():
    import nqes
    hip = Hip(nqes)
    yield nqes, hip
    hip.free()


def noalloc_case(nqes, hip, N, M, K, C, m, cplx, seed):
    print(f"case noalloc N={N} M={M} K={K} C={C} m={m} cplx={cplx}", flush=True)
    J = load_J(N)
    W = random_weights(N, M, K, seed=seed, scale=0.1, complex_=cplx)
    S = C * m
    rng = np.random.default_rng(seed)
    with nqes.Solver(N, M, K, 0.595, J) as s:
        P = s.P
        s.sample(W, seed, C, m, 8, 0.9)
        O = s.O_rows(0, S)
        mom = s.moments()
        v = rng.normal(size=P) + (1j * rng.normal(size=P) if cplx else 0.0)
        ref = s.apply_S(np.asarray(v, dtype=np.complex128))
    if cplx:
        ldo = P + 3  # a row stride past P (padding columns hold zeros)
        Oh = np.zeros((S, ldo), dtype=np.complex128)
        Oh[:, :P] = O
        conv = lambda x: np.ascontiguousarray(x, dtype=np.complex128)
    else:
        assert not np.any(O.imag) and not np.any(mom["o"].imag) and not np.any(mom["skk"].imag)
        ldo = P + (P & 1) + 2
        Oh = np.zeros((S, ldo))
        Oh[:, :P] = O.real
        conv = lambda x: np.ascontiguousarray(np.real(x), dtype=np.float64)
    Od, od, skkd, vd = hip.put(Oh), hip.put(conv(mom["o"])), hip.put(conv(mom["skk"])), hip.put(conv(v))
    like = conv(np.full(P, np.nan))
    outd = hip.put(like)  # NaN-filled: every entry must be written
    stream = ctypes.c_void_p()
    assert hip.rt.hipStreamCreate(ctypes.byref(stream)) == 0
    lib = nqes.load_library()
    st = lib.nqes_apply_S_rows_noalloc(Od, S, P, ldo, int(cplx), od, skkd, vd, outd, stream)
    assert st == 0, lib.nqes_last_error(None)
    assert hip.rt.hipStreamSynchronize(stream) == 0
    hip.rt.hipStreamDestroy(stream)
    res = hip.get(outd, like).astype(np.complex128)
    hip.free()
    err = rel_err(res, ref)
    print(f"noalloc vs nqes_apply_S: {err:.3e}", flush=True)
    assert err < 1e-12, err


@pytest.mark.parametrize("cplx", [True, False])
def test_noalloc_matches_apply_S(env, cplx):
    nqes, hip = env
    noalloc_case(nqes, hip, 16, 32, 2, 64, 16, cplx, seed=41)   # S = 1024, P = 1120
    noalloc_case(nqes, hip, 8, 16, 2, 16, 8, cplx, seed=43)     # smaller: S = 128, P = 304


def test_noalloc_refusals(env):
    nqes, hip = env
    lib = nqes.load_library()
    p = hip.put(np.zeros(16))
    # rows of 20000 complex entries: v would need 320 KB of LDS (refused before any launch)
    st = lib.nqes_apply_S_rows_noalloc(p, 4, 20000, 20000, 1, p, p, p, p, None)
    assert st == 1 and b"exceed" in lib.nqes_last_error(None)
    st = lib.nqes_apply_S_rows_noalloc(None, 4, 8, 8, 0, p, p, p, p, None)
    assert st == 1 and b"NULL" in lib.nqes_last_error(None)
    st = lib.nqes_apply_S_rows_noalloc(p, 4, 8, 7, 0, p, p, p, p, None)
    assert st == 1 and b"stride" in lib.nqes_last_error(None)
    hip.free()
