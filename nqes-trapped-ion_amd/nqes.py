"""Python host mirror of libnqes (include/nqes.h).

Mirrors the reference's SR-step seam (code/rbm.cc:564-720, ``monte_carlo``) and
its training loop (code/rbm.cc:1173-1197, ``run``) over the C ABI.  The HIP
library is the only compute path: if libnqes.so is missing or no GPU is
visible, constructing a :class:`Solver` raises.

Array conventions: complex arrays are numpy complex128; weights are one flat
vector of P = K*(N+M+N*M) entries, RBM k = [a(N), b(M), w(N*M, i-major)].
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libnqes.so")

PREC_FP64 = 0
PREC_MIXED = 1
MODEL_TRAPPEDION = 0  # ising-like: trapped-ion and the diagonal models (J given)
MODEL_HS = 1          # hs-like, delta = 1 (Haldane-Shastry with its J)
MODEL_XXZ = 2         # hs-like, delta = -1
ECANCELED = 6         # status of a solve ended by Server.cancel

_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


class NqesError(RuntimeError):
    pass


class Stats(ctypes.Structure):
    _fields_ = [
        ("minres_flag", ctypes.c_int),
        ("minres_iters", ctypes.c_uint64),
        ("accepts", ctypes.c_uint64),
        ("proposals", ctypes.c_uint64),
        ("sv_calls", ctypes.c_uint64),
        ("t_sample_ms", ctypes.c_double),
        ("t_moments_ms", ctypes.c_double),
        ("t_solve_ms", ctypes.c_double),
        ("t_sv_ms", ctypes.c_double),
        ("t_total_ms", ctypes.c_double),
        ("nonfinite", ctypes.c_int),
        ("ulp_ties", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Request(ctypes.Structure):
    """nqes_request: one whole solve (run() of rbm.cc:1173-1197)."""
    _fields_ = [
        ("N", ctypes.c_int), ("M", ctypes.c_int), ("K", ctypes.c_int),
        ("h", ctypes.c_double),
        ("J", ctypes.c_void_p),
        ("model", ctypes.c_int), ("precision", ctypes.c_int),
        ("W0", ctypes.c_void_p),
        ("seed", ctypes.c_uint64),
        ("C", ctypes.c_int), ("m", ctypes.c_int), ("B", ctypes.c_int),
        ("niters", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("lrate_init", ctypes.c_double), ("lrate_decay", ctypes.c_double), ("lrate_cutoff", ctypes.c_double),
        ("reglam_init", ctypes.c_double), ("reglam_decay", ctypes.c_double), ("reglam_cutoff", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("maxit", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("steps", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("K", ctypes.c_int),
        ("energies", ctypes.c_double * 4),
        ("t_submit_ms", ctypes.c_double),
        ("t_admit_ms", ctypes.c_double),
        ("t_done_ms", ctypes.c_double),
        ("t_device_ms", ctypes.c_double),
        ("bytes", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "energies"}
        d["energies"] = list(self.energies)[: self.K]
        return d


_lib = None

# every symbol include/nqes.h declares
EXPORTS = [
    "nqes_create", "nqes_destroy", "nqes_last_error", "nqes_problem_init", "nqes_sr_step",
    "nqes_sample", "nqes_get_samples", "nqes_get_O", "nqes_get_moments", "nqes_apply_S",
    "nqes_solve", "nqes_correlations", "nqes_required_bytes", "nqes_version",
    "nqes_set_dense_S", "nqes_enumerate", "nqes_get_ulp_ties", "nqes_get_sampler_launch",
    "nqes_apply_S_rows_noalloc",
    "nqes_server_create", "nqes_server_destroy", "nqes_server_last_error", "nqes_server_submit",
    "nqes_server_poll", "nqes_server_progress", "nqes_server_wait", "nqes_server_cancel",
]


def load_library(path=LIB_PATH):
    """Loads libnqes.so and declares the C signatures (no device call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NqesError(f"{path} not built: run nqes-trapped-ion_amd/build.py")
    L = ctypes.CDLL(path)
    c_int, c_u64, c_i64, c_dbl = ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_double
    L.nqes_create.restype = _vp
    L.nqes_create.argtypes = [c_int]
    L.nqes_destroy.argtypes = [_vp]
    L.nqes_last_error.restype = ctypes.c_char_p
    L.nqes_last_error.argtypes = [_vp]
    L.nqes_problem_init.restype = c_int
    L.nqes_problem_init.argtypes = [_vp, c_int, c_int, c_int, c_dbl, _dp, c_int, c_int]
    L.nqes_sr_step.restype = c_int
    L.nqes_sr_step.argtypes = [_vp, _cp, c_u64, c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_u64, _cp,
                               ctypes.POINTER(Stats)]
    L.nqes_sample.restype = c_int
    L.nqes_sample.argtypes = [_vp, _cp, c_u64, c_int, c_int, c_int, c_dbl]
    L.nqes_get_samples.restype = c_int
    L.nqes_get_samples.argtypes = [_vp, _vp, _vp, _vp]
    L.nqes_get_O.restype = c_int
    L.nqes_get_O.argtypes = [_vp, c_i64, c_i64, _cp]
    L.nqes_apply_S_rows_noalloc.restype = c_int
    L.nqes_apply_S_rows_noalloc.argtypes = [_vp, c_i64, c_i64, c_i64, c_int, _vp, _vp, _vp, _vp, _vp]
    L.nqes_get_ulp_ties.restype = c_int
    L.nqes_get_ulp_ties.argtypes = [_vp, _vp, ctypes.POINTER(c_u64)]
    L.nqes_get_sampler_launch.restype = c_int
    L.nqes_get_sampler_launch.argtypes = [_vp] + [ctypes.POINTER(c_int)] * 4
    L.nqes_get_moments.restype = c_int
    L.nqes_get_moments.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp]
    L.nqes_apply_S.restype = c_int
    L.nqes_apply_S.argtypes = [_vp, _cp, _cp]
    L.nqes_solve.restype = c_int
    L.nqes_solve.argtypes = [_vp, c_dbl, c_u64, _cp, ctypes.POINTER(c_int), ctypes.POINTER(c_u64)]
    L.nqes_correlations.restype = c_int
    L.nqes_correlations.argtypes = [_vp, _cp, c_u64, c_int, c_int, c_int, _cp, _cp, _cp, _cp]
    L.nqes_required_bytes.restype = c_i64
    L.nqes_required_bytes.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int]
    L.nqes_version.restype = ctypes.c_char_p
    L.nqes_version.argtypes = []
    L.nqes_set_dense_S.restype = c_int
    L.nqes_set_dense_S.argtypes = [_vp, c_int]
    L.nqes_enumerate.restype = c_int
    L.nqes_enumerate.argtypes = [_vp, _cp, c_i64, c_i64, _vp, _vp, _vp]
    L.nqes_server_create.restype = _vp
    L.nqes_server_create.argtypes = [c_int, c_int, c_i64]
    L.nqes_server_destroy.argtypes = [_vp]
    L.nqes_server_last_error.restype = ctypes.c_char_p
    L.nqes_server_last_error.argtypes = [_vp]
    L.nqes_server_submit.restype = c_int
    L.nqes_server_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(c_u64)]
    L.nqes_server_poll.restype = c_int
    L.nqes_server_poll.argtypes = [_vp, c_u64, ctypes.POINTER(c_int)]
    L.nqes_server_progress.restype = c_int
    L.nqes_server_progress.argtypes = [_vp, c_u64, ctypes.POINTER(c_int)]
    L.nqes_server_wait.restype = c_int
    L.nqes_server_wait.argtypes = [_vp, c_u64, _vp, ctypes.POINTER(Result)]
    L.nqes_server_cancel.restype = c_int
    L.nqes_server_cancel.argtypes = [_vp, c_u64]
    _lib = L
    return L


def nparams(N, M, K):
    return K * (N + M + N * M)


def nwords(N):
    return (N + 31) // 32


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


class Solver:
    """One device context with one problem (N, M, K, h, J) loaded.

    ``monte_carlo`` is the reference's SR step (rbm.cc:564-567): it updates
    ``W`` in place and returns the K x K sample mean of H_loc.
    """

    def __init__(self, N, M, K, h, J, precision=PREC_FP64, device=0, model=MODEL_TRAPPEDION):
        self.lib = load_library()
        self.ctx = self.lib.nqes_create(int(device))
        if not self.ctx:
            raise NqesError(f"nqes_create({device}) failed: no HIP device")
        self.N, self.M, self.K, self.h = int(N), int(M), int(K), float(h)
        self.P = nparams(N, M, K)
        self.NW = nwords(N)
        J = np.ascontiguousarray(J, dtype=np.float64)
        if J.shape != (N, N):
            raise ValueError(f"J must be {N}x{N}")
        self._check(self.lib.nqes_problem_init(self.ctx, self.N, self.M, self.K, self.h, J.reshape(-1),
                                               int(model), int(precision)))
        self.C = self.m = None

    def _check(self, st):
        if st != 0:
            msg = self.lib.nqes_last_error(self.ctx).decode()
            raise NqesError(f"nqes status {st}: {msg}")

    def close(self):
        if self.ctx:
            self.lib.nqes_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_dense_S(self, dense=True):
        """S as a dense P x P matrix (the reference's rbm-N* executables) instead of
        the matrix-free operator; FP64 only."""
        self._check(self.lib.nqes_set_dense_S(self.ctx, int(bool(dense))))

    def enumerate(self, W, cfg0=0, ncfg=None):
        """Phi (ncfg, K, K), local-energy numerator res (H_loc = Phi^+ res) and tanh
        (ncfg, K, K, M) of the configurations [cfg0, cfg0 + ncfg) (N K < 32)."""
        N, M, K = self.N, self.M, self.K
        if ncfg is None:
            ncfg = (1 << (N * K)) - cfg0
        W = np.ascontiguousarray(W, dtype=np.complex128)
        phi = np.zeros((ncfg, K, K), dtype=np.complex128)
        res = np.zeros((ncfg, K, K), dtype=np.complex128)
        tanh = np.zeros((ncfg, K, K, M), dtype=np.complex128)
        self._check(self.lib.nqes_enumerate(self.ctx, W, int(cfg0), int(ncfg), _ptr(phi), _ptr(res), _ptr(tanh)))
        return phi, res, tanh

    # ---------------------------------------------------------------- SR step
    def monte_carlo(self, W, seed, C, m, B, lr, reglam, rtol, maxit=2**64 - 1, stats=False):
        if W.dtype != np.complex128 or not W.flags.c_contiguous or W.shape != (self.P,):
            raise ValueError("W must be a contiguous complex128 vector of length P")
        emat = np.zeros(self.K * self.K, dtype=np.complex128)
        st = Stats()
        self._check(self.lib.nqes_sr_step(self.ctx, W, int(seed) % (1 << 64), int(C), int(m), int(B), float(lr),
                                          float(reglam), float(rtol), int(maxit), emat, ctypes.byref(st)))
        self.C, self.m = C, m
        emat = emat.reshape(self.K, self.K)
        return (emat, st.as_dict()) if stats else emat

    # ---------------------------------------------------------------- parts
    def sample(self, W, seed, C, m, B, reglam):
        W = np.ascontiguousarray(W, dtype=np.complex128)
        self._check(self.lib.nqes_sample(self.ctx, W, int(seed) % (1 << 64), int(C), int(m), int(B),
                                         float(reglam)))
        self.C, self.m = C, m

    def samples(self):
        S = self.C * self.m
        spins = np.zeros((S, self.K, self.NW), dtype=np.uint32)
        H = np.zeros((S, self.K, self.K), dtype=np.complex128)
        acc = np.zeros(self.C, dtype=np.uint64)
        self._check(self.lib.nqes_get_samples(self.ctx, _ptr(spins), _ptr(H), _ptr(acc)))
        return spins, H, acc

    def ulp_ties(self):
        """(per-chain, total) near-tie accept decisions of the last sample set (FP64)."""
        per = np.zeros(self.C, dtype=np.uint64)
        tot = ctypes.c_uint64(0)
        self._check(self.lib.nqes_get_ulp_ties(self.ctx, _ptr(per), ctypes.byref(tot)))
        return per, tot.value

    def sampler_launch(self):
        """Launch shape of the last mixed-precision sample set: {chains_per_block,
        tables_in_lds, samples_per_block, tuning} (tuning 1: a tuner trial, 0: the tuned
        choice, -1: tuner off)."""
        v = [ctypes.c_int(0) for _ in range(4)]
        self._check(self.lib.nqes_get_sampler_launch(self.ctx, *[ctypes.byref(x) for x in v]))
        return dict(zip(("chains_per_block", "tables_in_lds", "samples_per_block", "tuning"), (x.value for x in v)))

    def O_rows(self, row0, nrows):
        out = np.zeros((nrows, self.P), dtype=np.complex128)
        self._check(self.lib.nqes_get_O(self.ctx, int(row0), int(nrows), out.reshape(-1)))
        return out

    def moments(self):
        o, skk, f = (np.zeros(self.P, dtype=np.complex128) for _ in range(3))
        eloc = np.zeros(1, dtype=np.complex128)
        emat = np.zeros(self.K * self.K, dtype=np.complex128)
        self._check(self.lib.nqes_get_moments(self.ctx, _ptr(o), _ptr(skk), _ptr(f), _ptr(eloc), _ptr(emat)))
        return dict(o=o, skk=skk, f=f, eloc=eloc[0], emat=emat.reshape(self.K, self.K))

    def apply_S(self, v):
        v = np.ascontiguousarray(v, dtype=np.complex128)
        out = np.zeros(self.P, dtype=np.complex128)
        self._check(self.lib.nqes_apply_S(self.ctx, v, out))
        return out

    def correlations(self, W, seed, C, m, B):
        """Sample means of Phi^+ fcn(idx) (rbm.cc:730-829): x, z of shape (N, K, K),
        xx, zz of shape (N, N, K, K)."""
        N, K = self.N, self.K
        W = np.ascontiguousarray(W, dtype=np.complex128)
        x = np.zeros(N * K * K, dtype=np.complex128)
        z = np.zeros(N * K * K, dtype=np.complex128)
        xx = np.zeros(N * N * K * K, dtype=np.complex128)
        zz = np.zeros(N * N * K * K, dtype=np.complex128)
        self._check(self.lib.nqes_correlations(self.ctx, W, int(seed) % (1 << 64), int(C), int(m), int(B),
                                               x, z, xx, zz))
        return x.reshape(N, K, K), z.reshape(N, K, K), xx.reshape(N, N, K, K), zz.reshape(N, N, K, K)

    def solve(self, rtol, maxit=2**64 - 1):
        x = np.zeros(self.P, dtype=np.complex128)
        flag = ctypes.c_int(0)
        it = ctypes.c_uint64(0)
        self._check(self.lib.nqes_solve(self.ctx, float(rtol), int(maxit), x, ctypes.byref(flag), ctypes.byref(it)))
        return x, flag.value, it.value


def project_correlations(cor, emat):
    """rbm.cc:764-769 and 1198-1216: res[idx, k] = Re diag(V^-1 cor[idx] V)[k] with V
    the eigenvectors of e_mat, columns ordered by ascending real eigenvalue."""
    evals, V = np.linalg.eig(emat)
    order = np.argsort(evals.real, kind="stable")
    Vi = np.linalg.inv(V)
    flat = cor.reshape(-1, cor.shape[-2], cor.shape[-1])
    res = np.einsum("ij,njk,ki->ni", Vi, flat, V).real
    return res[:, order].reshape(cor.shape[:-2] + (cor.shape[-1],))


def run(solver, W, seed, C, m, B, niters, tol=1e-6, lrate0=0.1, lratef=0.99, lrate_cutoff=0.01,
        reglam0=0.01, reglam_decay=0.9, reglam_cutoff=0.001, rtol=1e-10, maxit=2**64 - 1, log=None):
    """The reference's SR loop (rbm.cc:1173-1197): schedules, seeds seed + t*C,
    sorted real eigenvalues of e_mat/N, convergence when ||evals - best|| < tol.
    Returns (energies per step, converged step or None, final e_mat/N)."""
    N = solver.N
    reglam, lrate = reglam0, lrate0
    best = None
    hist = []
    emat = None
    for it in range(niters):
        emat = solver.monte_carlo(W, seed + it * C, C, m, B, max(lrate, lrate_cutoff), max(reglam, reglam_cutoff),
                                  rtol, maxit) / N
        reglam *= reglam_decay
        lrate *= lratef
        ev = np.sort(np.linalg.eigvals(emat).real)
        hist.append(ev)
        if log:
            log(ev)
        if it and np.linalg.norm(ev - best) < tol:
            return hist, it + 1, emat
        if not it or ev.sum() < best.sum():
            best = ev
    return hist, None, emat


class Server:
    """Continuous batching of whole solves on one device (nqes_server_*).

    ``submit`` copies the request (J, initial weights) and returns an id; solves run
    concurrently, each on its own HIP stream, up to ``max_slots`` at a time and
    within ``budget`` device bytes (default: 90 % of the free memory).  ``wait``
    returns (final weights, result dict) and raises on a failed solve.
    """

    def __init__(self, max_slots=4, budget=0, device=0):
        self.lib = load_library()
        self.srv = self.lib.nqes_server_create(int(device), int(max_slots), int(budget))
        if not self.srv:
            raise NqesError(f"nqes_server_create({device}, {max_slots}) failed: no HIP device or bad arguments")
        self._P = {}

    def _err(self):
        return self.lib.nqes_server_last_error(self.srv).decode()

    def submit(self, N, M, K, h, J, W0, seed, C, m, B, niters, precision=PREC_MIXED, model=MODEL_TRAPPEDION,
               tol=1e-6, lrate0=0.1, lratef=0.99, lrate_cutoff=0.01, reglam0=0.01, reglam_decay=0.9,
               reglam_cutoff=0.001, rtol=1e-10, maxit=2**64 - 1):
        J = np.ascontiguousarray(J, dtype=np.float64)
        W0 = np.ascontiguousarray(W0, dtype=np.complex128)
        if J.shape != (N, N) or W0.shape != (nparams(N, M, K),):
            raise ValueError("J must be N x N and W0 of length P")
        q = Request(int(N), int(M), int(K), float(h), J.ctypes.data, int(model), int(precision), W0.ctypes.data,
                    int(seed) % (1 << 64), int(C), int(m), int(B), int(niters), float(tol), float(lrate0),
                    float(lratef), float(lrate_cutoff), float(reglam0), float(reglam_decay), float(reglam_cutoff),
                    float(rtol), int(maxit))
        rid = ctypes.c_uint64(0)
        st = self.lib.nqes_server_submit(self.srv, ctypes.byref(q), ctypes.byref(rid))
        if st != 0:
            raise NqesError(f"nqes_server_submit status {st}: {self._err()}")
        self._P[rid.value] = nparams(N, M, K)
        return rid.value

    def done(self, rid):
        d = ctypes.c_int(0)
        st = self.lib.nqes_server_poll(self.srv, int(rid), ctypes.byref(d))
        if st != 0:
            raise NqesError(f"nqes_server_poll status {st}: {self._err()}")
        return bool(d.value)

    def progress(self, rid):
        """SR steps solve `rid` has completed so far (0 while queued)."""
        n = ctypes.c_int(0)
        st = self.lib.nqes_server_progress(self.srv, int(rid), ctypes.byref(n))
        if st != 0:
            raise NqesError(f"nqes_server_progress status {st}: {self._err()}")
        return n.value

    def wait(self, rid, raise_on_error=True):
        """(final weights, result dict).  A failed solve raises, or with
        raise_on_error=False returns its result with the status and message."""
        W = np.zeros(self._P.get(rid, 0), dtype=np.complex128)
        res = Result()
        st = self.lib.nqes_server_wait(self.srv, int(rid), W.ctypes.data if W.size else None, ctypes.byref(res))
        if st == 0 or res.t_done_ms > 0:
            self._P.pop(rid, None)
        if st != 0 and raise_on_error:
            raise NqesError(f"solve {rid} failed with status {st}: {self._err()}")
        d = res.as_dict()
        if st != 0:
            d["error"] = self._err()
        return W, d

    def cancel(self, rid):
        """Cancels solve rid: a queued request ends at once, a running one after its
        current SR step (status ECANCELED); a finished one is left as it is."""
        st = self.lib.nqes_server_cancel(self.srv, int(rid))
        if st != 0:
            raise NqesError(f"nqes_server_cancel status {st}: {self._err()}")

    def close(self):
        if self.srv:
            self.lib.nqes_server_destroy(self.srv)
            self.srv = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
