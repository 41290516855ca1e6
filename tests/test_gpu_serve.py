"""Solve server: continuous batching of whole solves on one device (nqes_server_*).

* A request gives the same final weights, energies and step count, bit for bit, when
  it runs alone (one slot) and when it runs in a mix of concurrent solves that arrive
  while others are on the device (different sizes, K, precisions and real/complex
  weights).
* The server's solve is the reference's run() loop (rbm.cc:1173-1197): its final
  weights equal, bit for bit, those of the Python mirror's loop (nqes.run) over a
  Solver on the same request, and its energies agree to 1e-12 (numpy's eigvals vs
  the library's QR).
* Admission by device bytes: with a budget that holds one solve at a time, solves
  run one after another in arrival order; a request larger than the budget is
  refused at submit; invalid sizes are refused.
"""
import time

import numpy as np
import pytest

import nqes
from helpers import load_J, random_weights, synthetic_J

pytestmark = pytest.mark.gpu

PER_SOLVE = 64 << 20  # the server's fixed per-solve allowance

# (N, M, K, precision, complex weights, C, m, B, niters, seed)
MIX = [
    (8, 16, 2, nqes.PREC_MIXED, False, 64, 32, 16, 6, 11),
    (12, 24, 3, nqes.PREC_FP64, True, 32, 16, 8, 4, 12),
    (16, 16, 1, nqes.PREC_MIXED, True, 128, 32, 16, 5, 13),
    (32, 64, 2, nqes.PREC_MIXED, False, 256, 64, 16, 3, 14),
    (10, 10, 4, nqes.PREC_FP64, False, 32, 16, 8, 3, 15),
    (20, 40, 3, nqes.PREC_MIXED, True, 64, 32, 8, 4, 16),
]

SCHED = dict(tol=1e-6, lrate0=0.1, lratef=0.99, lrate_cutoff=0.01, reglam0=40.0, reglam_decay=0.9,
             reglam_cutoff=1.0, rtol=1e-9)


def request(case):
    N, M, K, prec, cplx, C, m, B, niters, seed = case
    J = load_J(N) if N in (8, 16, 32) else synthetic_J(N)
    W0 = random_weights(N, M, K, seed, scale=0.05 if cplx else 0.01, complex_=cplx)
    h = 0.3 + 0.05 * seed % 1.0
    return dict(N=N, M=M, K=K, h=h, J=J, W0=W0, seed=1000 * seed, C=C, m=m, B=B, niters=niters, precision=prec)


def run_alone(q):
    with nqes.Server(max_slots=1) as srv:
        rid = srv.submit(**q, **SCHED)
        return srv.wait(rid)


@pytest.fixture(scope="module")
def alone():
    out = []
    for case in MIX:
        print("alone:", case, flush=True)
        out.append(run_alone(request(case)))
    return out


def test_same_bits_alone_and_in_mix(alone):
    print("mix of", len(MIX), "solves, 4 slots, staggered arrivals", flush=True)
    with nqes.Server(max_slots=4) as srv:
        ids = []
        for case in MIX:
            ids.append(srv.submit(**request(case), **SCHED))
            time.sleep(0.005)  # staggered, yet well inside one solve however fast the kernels
        seen = [srv.progress(i) for i in ids]
        while not all(srv.done(i) for i in ids):
            now = [srv.progress(i) for i in ids]
            assert all(b >= a for a, b in zip(seen, now))  # steps completed only grow
            seen = now
            time.sleep(0.01)
        final = [srv.progress(i) for i in ids]
        res = [srv.wait(i) for i in ids]
    assert final == [r["steps"] for _, r in res]  # a finished solve reports all its steps
    overlap = 0
    for case, (Wa, ra), (Wm, rm) in zip(MIX, alone, res):
        assert np.array_equal(Wa, Wm), case
        assert ra["energies"] == rm["energies"], case
        assert ra["steps"] == rm["steps"] and ra["converged"] == rm["converged"], case
        assert rm["status"] == 0 and rm["steps"] >= 1
    # later arrivals were admitted while earlier solves were still running
    for a in range(len(res)):
        for b in range(a + 1, len(res)):
            ra, rb = res[a][1], res[b][1]
            if rb["t_admit_ms"] < ra["t_done_ms"] and ra["t_admit_ms"] < rb["t_done_ms"]:
                overlap += 1
    assert overlap >= 3


@pytest.mark.parametrize("idx", [0, 1, 5])
def test_server_equals_python_loop(alone, idx):
    case = MIX[idx]
    q = request(case)
    print("python loop:", case, flush=True)
    W = q["W0"].copy()
    with nqes.Solver(q["N"], q["M"], q["K"], q["h"], q["J"], precision=q["precision"]) as s:
        hist, conv, _ = nqes.run(s, W, q["seed"], q["C"], q["m"], q["B"], q["niters"], tol=SCHED["tol"],
                                 lrate0=SCHED["lrate0"], lratef=SCHED["lratef"],
                                 lrate_cutoff=SCHED["lrate_cutoff"], reglam0=SCHED["reglam0"],
                                 reglam_decay=SCHED["reglam_decay"], reglam_cutoff=SCHED["reglam_cutoff"],
                                 rtol=SCHED["rtol"])
    Ws, rs = alone[idx]
    assert np.array_equal(W, Ws)
    assert rs["steps"] == len(hist)
    assert rs["converged"] == (conv is not None)
    np.testing.assert_allclose(rs["energies"], hist[-1], rtol=1e-12, atol=1e-12)


def test_admission_by_device_bytes(alone):
    lib = nqes.load_library()
    cases = [MIX[0], MIX[2], MIX[0]]  # similar footprints, so a budget can hold one but never two
    need = [lib.nqes_required_bytes(c[0], c[1], c[2], c[5], c[6], c[3]) + PER_SOLVE for c in cases]
    budget = int(max(need) * 1.5)  # any two together exceed it
    assert all(a + b > budget for a in need for b in need)
    print("budget", budget, "needs", need, flush=True)
    with nqes.Server(max_slots=4, budget=budget) as srv:
        ids = [srv.submit(**request(c), **SCHED) for c in cases]
        res = [srv.wait(i) for i in ids]
    for c, (W, r) in zip(cases, res):
        Wa, ra = alone[MIX.index(c)]
        assert np.array_equal(W, Wa) and r["energies"] == ra["energies"]
        assert r["bytes"] == need[cases.index(c)]
    for a, b in zip(res, res[1:]):  # one at a time, arrival order
        assert b[1]["t_admit_ms"] >= a[1]["t_done_ms"]


def test_refusals():
    q = request(MIX[0])
    with nqes.Server(max_slots=1, budget=PER_SOLVE) as srv:
        with pytest.raises(nqes.NqesError, match="status 2"):
            srv.submit(**q, **SCHED)
        bad = dict(q, M=12, W0=random_weights(8, 12, 2, 1))  # M % N != 0
        with pytest.raises(nqes.NqesError, match="status 1"):
            srv.submit(**bad, **SCHED)


def scan_and_serving_mix():
    """Requests drawn as the batched field scan (N in {16, 24, 32, 48}, M = 2N, K in
    {2, 3, 4}, h ~ U[0.1, 1.5], J exponent 1) and the serving stream (N even in
    16..64, M in {N, 2N}, K in 1..4, h ~ U[0.2, 1.2], J exponent ~ U[0.5, 1.5],
    including the hs and xxz pair-flip models), at small chain counts; complex
    init_rbm-style weights on the mixed path, real small ones on some FP64 requests."""
    rng = np.random.default_rng(8)
    out = []
    for r in range(8):
        if r < 4:
            N = int(rng.choice([16, 24, 32, 48]))
            M, K = 2 * N, int(rng.choice([2, 3, 4]))
            h, alpha, model = float(rng.uniform(0.1, 1.5)), 1.0, nqes.MODEL_TRAPPEDION
        else:
            N = int(rng.choice(np.arange(16, 65, 2)))
            M, K = N * int(rng.choice([1, 2])), int(rng.integers(1, 5))
            h, alpha = float(rng.uniform(0.2, 1.2)), float(rng.uniform(0.5, 1.5))
            model = [nqes.MODEL_TRAPPEDION, nqes.MODEL_HS, nqes.MODEL_XXZ, nqes.MODEL_TRAPPEDION][r - 4]
        prec = nqes.PREC_FP64 if r in (2, 6) else nqes.PREC_MIXED
        cplx = prec == nqes.PREC_MIXED or r == 6
        J = synthetic_J(N, alpha)
        if model != nqes.MODEL_TRAPPEDION:
            J = np.tril(J, -1) * (-1.0 if model == nqes.MODEL_XXZ else 1.0)
        W0 = random_weights(N, M, K, 500 + r, scale=0.3 if cplx else 0.01, complex_=cplx)
        out.append(dict(N=N, M=M, K=K, h=h, J=J, W0=W0, seed=31 * r + 7, C=32 if prec == nqes.PREC_FP64 else 128,
                        m=16, B=8, niters=3, precision=prec, model=model))
    return out


SCHED_PUB = dict(tol=1e-4, lrate0=0.1, lratef=0.99, lrate_cutoff=0.01, reglam0=1.0, reglam_decay=0.9,
                 reglam_cutoff=0.2, rtol=1e-9, maxit=3000)


def test_scan_and_serving_mix_alone_equals_mix():
    """Each request of the scan / serving mix gives the same bits alone (its own fresh
    server) and in the mix, where three slots keep their contexts across solves of
    different sizes, K, models and precisions (buffers grown by earlier solves reused)."""
    reqs = scan_and_serving_mix()
    alone = []
    for q in reqs:
        print("alone:", {k: q[k] for k in ("N", "M", "K", "model", "precision")}, flush=True)
        with nqes.Server(max_slots=1) as srv:
            alone.append(srv.wait(srv.submit(**q, **SCHED_PUB)))
    with nqes.Server(max_slots=3) as srv:
        ids = []
        for q in reqs:
            ids.append(srv.submit(**q, **SCHED_PUB))
            time.sleep(0.002)
        mix = [srv.wait(i) for i in ids]
    for q, (Wa, ra), (Wm, rm) in zip(reqs, alone, mix):
        assert ra["status"] == 0 and rm["status"] == 0, (ra, rm)
        assert np.array_equal(Wa, Wm), {k: q[k] for k in ("N", "M", "K", "model", "precision")}
        assert ra["energies"] == rm["energies"] and ra["steps"] == rm["steps"]


def test_pairflip_request_equals_python_loop():
    """hs requests are served (rbm.cc:1092-1119 through the server) and equal the
    Python loop over a Solver bit for bit."""
    N, M, K = 12, 24, 2
    J = np.tril(synthetic_J(N, 2.0), -1)
    W0 = random_weights(N, M, K, 77, scale=0.2, complex_=True)
    q = dict(N=N, M=M, K=K, h=0.0, J=J, W0=W0, seed=99, C=64, m=16, B=8, niters=3, precision=nqes.PREC_MIXED,
             model=nqes.MODEL_HS)
    with nqes.Server(max_slots=2) as srv:
        Ws, rs = srv.wait(srv.submit(**q, **SCHED_PUB))
    W = W0.copy()
    with nqes.Solver(N, M, K, 0.0, J, precision=nqes.PREC_MIXED, model=nqes.MODEL_HS) as s:
        hist, conv, _ = nqes.run(s, W, 99, 64, 16, 8, 3, tol=1e-4, lrate0=0.1, lratef=0.99, lrate_cutoff=0.01,
                                 reglam0=1.0, reglam_decay=0.9, reglam_cutoff=0.2, rtol=1e-9, maxit=3000)
    assert rs["status"] == 0 and np.array_equal(W, Ws) and rs["steps"] == len(hist)


def test_second_waiter_refused():
    """Each request id has one waiter: a second wait on an id that is being waited for
    (or already collected) is refused with NQES_EINVAL, never a use after free."""
    import threading
    q = request(MIX[3])
    with nqes.Server(max_slots=1) as srv:
        rid = srv.submit(**dict(q, niters=3), **SCHED)
        first = {}
        t = threading.Thread(target=lambda: first.update(r=srv.wait(rid)))
        t.start()
        time.sleep(0.05)
        with pytest.raises(nqes.NqesError, match="status 1"):
            srv.wait(rid)
        t.join()
        assert first["r"][1]["status"] == 0


def test_concurrent_launches_of_one_kernel_with_different_lds():
    """Two contexts on two threads sample with complex weights at N=28 and N=32 (K=2,
    M=2N): the same sweep and measurement kernels, each launched with more than 64 KB
    of dynamic LDS and a different amount per N.  Kernel attributes are shared by the
    whole process, so a per-launch maximum set by one thread could undercut the other
    thread's launch (an intermittent launch failure of concurrent solves); the library
    sets it once per kernel.  Every call succeeds and each thread's samples equal those
    of the same calls made alone."""
    import threading

    cases = [(28, 56), (32, 64)]
    reps = 30

    def run(N, M, out, errs):
        J = synthetic_J(N)
        W = random_weights(N, M, 2, 7 + N, scale=0.05, complex_=True)
        try:
            with nqes.Solver(N, M, 2, 0.5, J, precision=nqes.PREC_MIXED) as s:
                for r in range(reps):
                    s.sample(W, 100 * r + N, 256, 8, 4, 1.0)
                out.append(s.samples())
        except nqes.NqesError as e:
            errs.append(f"N={N}: {e}")

    alone = {}
    for N, M in cases:
        o, e = [], []
        run(N, M, o, e)
        assert not e, e
        alone[N] = o[0]
    print("alone done; now concurrent", flush=True)
    outs = {N: [] for N, _ in cases}
    errs = []
    th = [threading.Thread(target=run, args=(N, M, outs[N], errs)) for N, M in cases]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errs, errs
    for N, _ in cases:
        for a, b in zip(alone[N], outs[N][0]):
            assert np.array_equal(a, b), N


def test_cancel_queued_and_running():
    """nqes_server_cancel (a serving deadline): a queued request ends at once with status
    ECANCELED and no step run; a running one ends after its current SR step, with the
    weights of its completed steps, equal bit for bit to the same request run for that
    many steps; the server then serves the next request as before."""
    q = request(MIX[3])
    sched = dict(SCHED, tol=0.0)  # never converges: only the cancel ends the solve
    with nqes.Server(max_slots=1) as srv:
        r1 = srv.submit(**dict(q, niters=400), **sched)
        r2 = srv.submit(**dict(q, niters=400), **sched)
        t0 = time.time()
        while srv.progress(r1) < 2 and time.time() - t0 < 120:
            time.sleep(0.01)
        srv.cancel(r2)
        _, d2 = srv.wait(r2, raise_on_error=False)
        assert d2["status"] == nqes.ECANCELED and d2["steps"] == 0
        srv.cancel(r1)
        W1, d1 = srv.wait(r1, raise_on_error=False)
        assert d1["status"] == nqes.ECANCELED and 2 <= d1["steps"] < 400, d1
        Wn, dn = srv.wait(srv.submit(**dict(q, niters=d1["steps"]), **sched))
        assert dn["status"] == 0 and dn["steps"] == d1["steps"]
        assert np.array_equal(W1, Wn)
        assert np.array_equal(np.asarray(d1["energies"]), np.asarray(dn["energies"]))

