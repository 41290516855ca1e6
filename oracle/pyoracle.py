"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product never
imports this module.

Array conventions (shared with the C ABI in include/nqes.h):
  complex arrays are interleaved float64 (numpy complex128);
  weights W: P = K*L complex, RBM kk block [kk*L, (kk+1)*L) = a[N], b[M], w[N][M];
  spins: uint32 words, NW = ceil(N/32) per replica, bit i of replica k in
         word k*NW + i//32, bit i%32 (1 means sigma_i = -1, rbm.cc:141);
  tanh: [k][kk][j] (replica k, RBM kk, hidden unit j);
  Phi: K x K row-major [replica k][RBM kk].
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        c_int, c_long, c_dbl, c_u64 = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_uint64
        vp = ctypes.c_void_p
        L.orc_exp.restype = c_dbl
        L.orc_exp.argtypes = [c_dbl]
        L.orc_sincos_v.argtypes = [c_dbl, _dp]
        L.orc_coshsinh_v.argtypes = [c_dbl, _dp]
        for f in ("orc_cexp", "orc_ccosh", "orc_csinh", "orc_ctanh"):
            getattr(L, f).argtypes = [_dp, _dp]
        L.orc_cdiv.argtypes = [_dp, _dp, _dp]
        L.orc_tree_prod.argtypes = [_cp, c_int, _dp]
        L.orc_pcg64_init.argtypes = [c_u64, _u64p]
        L.orc_pcg64_next.restype = c_u64
        L.orc_pcg64_next.argtypes = [_u64p]
        L.orc_canonical.restype = c_dbl
        L.orc_canonical.argtypes = [c_u64]
        L.orc_det.argtypes = [c_int, _cp, _dp]
        L.orc_inv.restype = c_int
        L.orc_inv.argtypes = [c_int, _cp, _cp]
        L.orc_state_fcn.argtypes = [c_int, c_int, c_int, _cp, _up, _cp, _cp]
        L.orc_flip.argtypes = [c_int, c_int, c_int, _cp, c_int, c_int, _up, _cp, _cp]
        L.orc_ezz.restype = c_dbl
        L.orc_ezz.argtypes = [c_int, _dp, _up]
        L.orc_hloc.argtypes = [c_int, c_int, c_int, c_dbl, _dp, _cp, _up, _cp, _cp, _cp, _cp]
        L.orc_hloc_model.argtypes = [c_int] + L.orc_hloc.argtypes
        L.orc_deriv.argtypes = [c_int, c_int, c_int, _up, _cp, _cp, _cp, _cp]
        L.orc_run_chain.argtypes = [c_int, c_int, c_int, c_dbl, _dp, _cp, c_u64, c_int, c_int, c_int,
                                    vp, vp, vp, vp, vp, vp, ctypes.POINTER(c_long)]
        L.orc_run_chain_model.argtypes = [c_int] + L.orc_run_chain.argtypes
        L.orc_sample_chains_model.argtypes = [c_int, c_int, c_int, c_int, c_dbl, _dp, _cp, c_u64, c_int, c_int,
                                              c_int, vp, vp, vp, vp]
        L.orc_minresqlp_dense.restype = c_int
        L.orc_minresqlp_dense.argtypes = [c_long, _cp, _cp, _cp, c_dbl, c_u64, ctypes.POINTER(c_u64)]
        L.orc_sr_step.restype = c_int
        L.orc_sr_step.argtypes = [c_int, c_int, c_int, c_dbl, _dp, _cp, c_u64, c_int, c_int, c_int,
                                  c_dbl, c_dbl, c_dbl, c_u64, c_int, _cp, ctypes.POINTER(c_u64),
                                  ctypes.POINTER(c_long), vp, vp, vp, vp, vp]
        L.orc_sr_step_model.restype = c_int
        L.orc_sr_step_model.argtypes = [c_int] + L.orc_sr_step.argtypes
        L.orc_sv_from_O.restype = c_int
        L.orc_sv_from_O.argtypes = [c_long, c_long, _cp, _cp, _cp, _cp, _cp]
        L.orc_init_spins.argtypes = [c_int, c_int, c_u64, c_int, _up]
        L.orc_correlations.argtypes = [c_int, c_int, c_int, c_dbl, _dp, _cp, c_u64, c_int, c_int, c_int,
                                       _cp, _cp, _cp, _cp]
        L.orc_max_threads.restype = c_int
        L.orc_set_threads.argtypes = [c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- scalars
def exp(x):
    return lib().orc_exp(float(x))


def sincos(x):
    o = np.zeros(2)
    lib().orc_sincos_v(float(x), o)
    return o[0], o[1]


def coshsinh(x):
    o = np.zeros(2)
    lib().orc_coshsinh_v(float(x), o)
    return o[0], o[1]


def _cfun(name, z):
    i = np.array([z.real, z.imag], dtype=np.float64)
    o = np.zeros(2)
    getattr(lib(), name)(i, o)
    return complex(o[0], o[1])


def cexp(z):
    return _cfun("orc_cexp", complex(z))


def ccosh(z):
    return _cfun("orc_ccosh", complex(z))


def csinh(z):
    return _cfun("orc_csinh", complex(z))


def ctanh(z):
    return _cfun("orc_ctanh", complex(z))


def cdiv(a, b):
    a, b = complex(a), complex(b)
    o = np.zeros(2)
    lib().orc_cdiv(np.array([a.real, a.imag]), np.array([b.real, b.imag]), o)
    return complex(o[0], o[1])


def tree_prod(d):
    d = np.ascontiguousarray(d, dtype=np.complex128)
    o = np.zeros(2)
    lib().orc_tree_prod(d, len(d), o)
    return complex(o[0], o[1])


class PCG64:
    """pcg64(seed) of the reference (setseq_xsl_rr_128_64, default stream)."""

    def __init__(self, seed):
        self.st = np.zeros(2, dtype=np.uint64)
        lib().orc_pcg64_init(int(seed) % (1 << 64), self.st)

    def __call__(self):
        return int(lib().orc_pcg64_next(self.st))


def canonical(u):
    return lib().orc_canonical(int(u))


# ---------------------------------------------------------------- model pieces
def nwords(N):
    return (N + 31) // 32


def nparams(N, M, K):
    return K * (N + M + N * M)


def pack_spins(bits):
    """bits: (..., K, N) array of 0/1 -> (..., K, NW) uint32 words."""
    bits = np.asarray(bits, dtype=np.uint64)
    N = bits.shape[-1]
    NW = nwords(N)
    out = np.zeros(bits.shape[:-1] + (NW,), dtype=np.uint32)
    for i in range(N):
        out[..., i // 32] |= (bits[..., i] << np.uint64(i % 32)).astype(np.uint32)
    return out


def unpack_spins(words, N):
    words = np.asarray(words, dtype=np.uint32)
    bits = np.zeros(words.shape[:-1] + (N,), dtype=np.uint8)
    for i in range(N):
        bits[..., i] = (words[..., i // 32] >> np.uint32(i % 32)) & 1
    return bits


def init_spins(N, K, seed, c):
    out = np.zeros(K * nwords(N), dtype=np.uint32)
    lib().orc_init_spins(N, K, int(seed) % (1 << 64), c, out)
    return out.reshape(K, nwords(N))


def det(m):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    o = np.zeros(2)
    lib().orc_det(m.shape[0], m.reshape(-1), o)
    return complex(o[0], o[1])


def inv(m):
    m = np.ascontiguousarray(m, dtype=np.complex128)
    out = np.zeros_like(m)
    ok = lib().orc_inv(m.shape[0], m.reshape(-1), out.reshape(-1))
    return out, bool(ok)


def state_fcn(N, M, K, W, spins):
    W = np.ascontiguousarray(W, dtype=np.complex128)
    spins = np.ascontiguousarray(spins, dtype=np.uint32).reshape(-1)
    phi = np.zeros(K * K, dtype=np.complex128)
    tanh = np.zeros(K * K * M, dtype=np.complex128)
    lib().orc_state_fcn(N, M, K, W, spins, phi, tanh)
    return phi.reshape(K, K), tanh.reshape(K, K, M)


def flip(N, M, K, W, k, i, spins, phi, tanh):
    phi = np.ascontiguousarray(phi, dtype=np.complex128).copy()
    tanh = np.ascontiguousarray(tanh, dtype=np.complex128).copy()
    lib().orc_flip(N, M, K, np.ascontiguousarray(W), k, i,
                   np.ascontiguousarray(spins, dtype=np.uint32).reshape(-1),
                   phi.reshape(-1), tanh.reshape(-1))
    return phi, tanh


def ezz(N, J, spins_k):
    return lib().orc_ezz(N, np.ascontiguousarray(J, dtype=np.float64).reshape(-1),
                         np.ascontiguousarray(spins_k, dtype=np.uint32).reshape(-1))


MODEL_TRAPPEDION, MODEL_HS, MODEL_XXZ = 0, 1, 2  # ising-like; hs-like with delta = 1, -1


def hloc(N, M, K, h, J, W, spins, phi, phi_inv, tanh, model=MODEL_TRAPPEDION):
    H = np.zeros(K * K, dtype=np.complex128)
    lib().orc_hloc_model(int(model), N, M, K, float(h), np.ascontiguousarray(J, dtype=np.float64).reshape(-1),
                   np.ascontiguousarray(W), np.ascontiguousarray(spins, dtype=np.uint32).reshape(-1),
                   np.ascontiguousarray(phi).reshape(-1), np.ascontiguousarray(phi_inv).reshape(-1),
                   np.ascontiguousarray(tanh).reshape(-1), H)
    return H.reshape(K, K)


def deriv(N, M, K, spins, phi, phi_inv, tanh):
    O = np.zeros(nparams(N, M, K), dtype=np.complex128)
    lib().orc_deriv(N, M, K, np.ascontiguousarray(spins, dtype=np.uint32).reshape(-1),
                    np.ascontiguousarray(phi).reshape(-1), np.ascontiguousarray(phi_inv).reshape(-1),
                    np.ascontiguousarray(tanh).reshape(-1), O)
    return O


def run_chain(N, M, K, h, J, W, seed, c, m, B, want_O=False, want_state=False, model=MODEL_TRAPPEDION):
    """One chain of an SR step: returns dict of per-sample spins, H, [O], [state]."""
    NW = nwords(N)
    P = nparams(N, M, K)
    spins = np.zeros((m, K, NW), dtype=np.uint32)
    H = np.zeros((m, K, K), dtype=np.complex128)
    O = np.zeros((m, P), dtype=np.complex128) if want_O else None
    phi = np.zeros((m, K, K), dtype=np.complex128) if want_state else None
    phi_inv = np.zeros((m, K, K), dtype=np.complex128) if want_state else None
    tanh = np.zeros((m, K, K, M), dtype=np.complex128) if want_state else None
    acc = ctypes.c_long(0)
    lib().orc_run_chain_model(int(model), N, M, K, float(h), np.ascontiguousarray(J, dtype=np.float64).reshape(-1),
                        np.ascontiguousarray(W, dtype=np.complex128), int(seed) % (1 << 64), c, m, B,
                        _ptr(spins), _ptr(H), _ptr(O), _ptr(phi), _ptr(phi_inv), _ptr(tanh),
                        ctypes.byref(acc))
    out = dict(spins=spins, H=H, accepts=acc.value)
    if want_O:
        out["O"] = O
    if want_state:
        out.update(phi=phi, phi_inv=phi_inv, tanh=tanh)
    return out


def sample_chains(N, M, K, h, J, W, seed, C, m, B, want_O=False, model=MODEL_TRAPPEDION):
    """Chains 0..C-1 of an SR step in parallel: dict of spins (S, K, NW), H (S, K, K),
    accepts (C,) and, with want_O, O (S, P); sample s = c * m + t."""
    NW = nwords(N)
    P = nparams(N, M, K)
    S = C * m
    spins = np.zeros((S, K, NW), dtype=np.uint32)
    H = np.zeros((S, K, K), dtype=np.complex128)
    O = np.zeros((S, P), dtype=np.complex128) if want_O else None
    acc = np.zeros(C, dtype=np.int64)
    lib().orc_sample_chains_model(int(model), N, M, K, float(h), np.ascontiguousarray(J, dtype=np.float64).reshape(-1),
                                  np.ascontiguousarray(W, dtype=np.complex128), int(seed) % (1 << 64), C, m, B,
                                  _ptr(spins), _ptr(H), _ptr(O), _ptr(acc))
    out = dict(spins=spins, H=H, accepts=acc.astype(np.uint64))
    if want_O:
        out["O"] = O
    return out


def minresqlp_dense(A, b, rtol, maxit=2**64 - 1):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b)
    it = ctypes.c_uint64(0)
    flag = lib().orc_minresqlp_dense(len(b), A.reshape(-1), b, x, float(rtol), int(maxit), ctypes.byref(it))
    return x, flag, it.value


def sr_step(N, M, K, h, J, W, seed, C, m, B, lr, reglam, rtol, maxit=2**64 - 1, store_O=False,
            want_details=False, model=MODEL_TRAPPEDION):
    """One SR step (rbm.cc monte_carlo).  Returns (W_new, emat, info)."""
    P = nparams(N, M, K)
    Wn = np.ascontiguousarray(W, dtype=np.complex128).copy()
    emat = np.zeros(K * K, dtype=np.complex128)
    it = ctypes.c_uint64(0)
    acc = ctypes.c_long(0)
    det_arrays = [np.zeros(P, dtype=np.complex128) for _ in range(4)] if want_details else [None] * 4
    eloc = np.zeros(1, dtype=np.complex128) if want_details else None
    flag = lib().orc_sr_step_model(int(model), N, M, K, float(h), np.ascontiguousarray(J, dtype=np.float64).reshape(-1),
                             Wn, int(seed) % (1 << 64), C, m, B, float(lr), float(reglam), float(rtol),
                             int(maxit), int(bool(store_O)), emat, ctypes.byref(it), ctypes.byref(acc),
                             *[_ptr(a) for a in det_arrays], _ptr(eloc))
    info = dict(flag=flag, iters=it.value, accepts=acc.value)
    if want_details:
        info.update(x=det_arrays[0], o=det_arrays[1], skk=det_arrays[2], f=det_arrays[3], eloc=eloc[0])
    return Wn, emat.reshape(K, K), info


def sv_from_O(O, o, skk, v):
    O = np.ascontiguousarray(O, dtype=np.complex128)
    S, P = O.shape
    out = np.zeros(P, dtype=np.complex128)
    lib().orc_sv_from_O(S, P, O.reshape(-1), np.ascontiguousarray(o, dtype=np.complex128),
                        np.ascontiguousarray(skk, dtype=np.complex128),
                        np.ascontiguousarray(v, dtype=np.complex128), out)
    return out


def correlations(N, M, K, h, J, W, seed, C, m, B):
    """Sample means of Phi^+ fcn(idx) for sigma^x_i, sigma^z_i, sigma^x_i sigma^x_j,
    sigma^z_i sigma^z_j (rbm.cc:730-829); arrays (N, K, K) and (N, N, K, K)."""
    x = np.zeros(N * K * K, dtype=np.complex128)
    z = np.zeros(N * K * K, dtype=np.complex128)
    xx = np.zeros(N * N * K * K, dtype=np.complex128)
    zz = np.zeros(N * N * K * K, dtype=np.complex128)
    lib().orc_correlations(N, M, K, float(h), np.ascontiguousarray(J, dtype=np.float64).reshape(-1),
                           np.ascontiguousarray(W, dtype=np.complex128), int(seed) % (1 << 64), C, m, B, x, z, xx, zz)
    return x.reshape(N, K, K), z.reshape(N, K, K), xx.reshape(N, N, K, K), zz.reshape(N, N, K, K)


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()
