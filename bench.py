"""Benchmark: SR steps of the multi-RBM VMC solver on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision mixed|fp64]
                    [--N 32 --M 64 --K 2 --C 4096 --m 256 --B 64] [--no-cpu]

Workload (default): trapped-ion chain N=32 spins, M=64 hidden units, K=2 states,
h=0.595, J = power law of the N=32 ion chain, 4096 chains x 256 samples (S=2^20)
with 64 burn-in sweeps, real weights N(0, 0.01^2), lr 0.1, reglam 40, MINRES-QLP
rtol 1e-9.  A "step" is one full SR step: sampling + H_loc + O, moments, the
MINRES-QLP solve with its matrix-free S*v passes, and the weight update.

Prints one JSON line.  value = MC samples per second over whole SR steps.

    python bench.py --workload scan     [--slots 8] [--solve-steps 5]
    python bench.py --workload serving  [--slots 8] [--solve-steps 5] [--requests 48]

run whole solves through the solve server (continuous batching on one device):
  scan: 64 solves submitted at once, h ~ U[0.1, 1.5], N in {16, 24, 32, 48}, M = 2N,
        K in {2, 3, 4}, J_ij = |i-j|^-1 normalised to max 1, 1024 chains x 256 samples;
  serving: Poisson arrivals (4 requests/s), N in {16, 18, ..., 64}, M in {N, 2N},
        K in {1..4}, h ~ U[0.2, 1.2], J exponent ~ U[0.5, 1.5], 1024 chains x 256 samples.
Both use the mixed path and cap every solve at --solve-steps SR steps (the convergence
test still applies).  value = solves per hour; p50/p99 = arrival -> result latency.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "nqes-trapped-ion_amd"))

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix, published
INT8_MFMA_PEAK_TOPS = 5000.0  # MI355X dense int8 matrix, published


def load_J(N):
    path = os.path.join(REPO, "tests", "golden", f"J_N{N}.npy")
    if os.path.exists(path):
        return np.load(path)
    i = np.arange(N)
    d = np.abs(i[:, None] - i[None, :]).astype(float)
    J = np.zeros((N, N))
    J[d > 0] = d[d > 0] ** -1.0
    return J / J.max()


def initial_weights(N, M, K, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=K * (N + M + N * M)) * 0.01).astype(np.complex128)


def profiled_traffic(args, real, mode):
    """HBM bytes per S*v (all its kernels) measured by rocprofv3 (FETCH_SIZE x 2 +
    WRITE_SIZE, copy-calibrated) for this exact workload and S*v form, from the
    summaries under profiles/; None when no summary covers it."""
    key = (f"traffic: precision={args.precision} sv={mode} real={int(real)} N={args.N} M={args.M} K={args.K} "
           f"C={args.C} m={args.m} bytes=")
    pdir = os.path.join(REPO, "profiles")
    for name in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                line = line.strip()
                if line.startswith(key):
                    return int(line[len(key):].split()[0])
    return None


def power_law_J(N, alpha):
    i = np.arange(N)
    d = np.abs(i[:, None] - i[None, :]).astype(float)
    J = np.zeros((N, N))
    J[d > 0] = d[d > 0] ** -alpha
    return J / J.max()


def solve_requests(args):
    """(arrival time s, request) pairs of the scan or serving workload, seeded."""
    rng = np.random.default_rng(3 if args.workload == "scan" else 4)
    reqs = []
    t = 0.0
    n = 64 if args.workload == "scan" else args.requests
    for r in range(n):
        if args.workload == "scan":
            N = int(rng.choice([16, 24, 32, 48]))
            M, K = 2 * N, int(rng.choice([2, 3, 4]))
            h = float(rng.uniform(0.1, 1.5))
            alpha = 1.0
        else:
            t += float(rng.exponential(1.0 / args.rate))
            N = int(rng.choice(np.arange(16, 65, 2)))
            M = N * int(rng.choice([1, 2]))
            K = int(rng.integers(1, 5))
            h = float(rng.uniform(0.2, 1.2))
            alpha = float(rng.uniform(0.5, 1.5))
        W0 = (np.random.default_rng(1000 + r).normal(size=K * (N + M + N * M)) * 0.01).astype(np.complex128)
        reqs.append((t if args.workload == "serving" else 0.0,
                     dict(N=N, M=M, K=K, h=h, J=power_law_J(N, alpha), W0=W0, seed=args.seed + 7919 * r,
                          C=1024, m=256, B=args.B, niters=args.solve_steps)))
    return reqs


def serve_workload(args):
    import nqes
    reqs = solve_requests(args)
    sched = dict(tol=1e-6, lrate0=0.1, lratef=0.99, lrate_cutoff=0.01, reglam0=40.0, reglam_decay=0.9,
                 reglam_cutoff=1.0, rtol=args.rtol)
    with nqes.Server(max_slots=args.slots) as srv:
        # warm-up: one small solve (module load, first allocations)
        w = reqs[0][1]
        srv.wait(srv.submit(**dict(w, niters=1, C=64, m=16), precision=nqes.PREC_MIXED, **sched))
        ids = []
        t0 = time.perf_counter()
        for t_arr, q in reqs:
            dt = t_arr - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(**q, precision=nqes.PREC_MIXED, **sched))
        res = []
        for k, i in enumerate(ids):
            res.append(srv.wait(i, raise_on_error=False)[1])
            if (k + 1) % 8 == 0:
                print(f"[bench] {k + 1}/{len(ids)} solves done, {time.perf_counter() - t0:.1f}s",
                      file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
    lat = np.array([r["t_done_ms"] - r["t_submit_ms"] for r in res])
    queue = np.array([r["t_admit_ms"] - r["t_submit_ms"] for r in res])
    steps = sum(r["steps"] for r in res)
    samples = sum(r["steps"] * q["C"] * q["m"] for (_, q), r in zip(reqs, res))
    name = "batched field scan" if args.workload == "scan" else "serving (Poisson arrivals, 4 req/s)"
    out = {
        "metric": "MC samples/s (whole SR step); converged solves/hour; p50/p99 solve latency",
        "value": len(res) * 3600.0 / wall,
        "unit": f"solves/hour ({args.solve_steps} SR steps max per solve)",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 1,
        "ms_per_step": wall * 1e3 / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (power-law J, random N(0,0.01^2) real weights, seeded)",
        "config": {"workload": f"{name}: {len(res)} solves, 1024 chains x 256 samples, B={args.B}, "
                               f"<= {args.solve_steps} SR steps each, {args.slots} slots, mixed precision",
                   "solves": len(res), "slots": args.slots, "solve_steps": args.solve_steps},
        "p50_latency_ms": float(np.percentile(lat, 50)),
        "p99_latency_ms": float(np.percentile(lat, 99)),
        "p50_queue_ms": float(np.percentile(queue, 50)),
        "wall_s": wall,
        "mc_samples_per_s": samples / wall,
        "device_busy_ms_sum": float(sum(r["t_device_ms"] for r in res)),
        "failed": {"count": int(sum(r["status"] != 0 for r in res)),
                   "note": "status 4 = non-finite energy or gradient: the SR schedule (lr 0.1, reglam 40 -> 1) "
                           "diverges for some K=1 requests at small h, identically on the FP64 and mixed paths"},
    }
    print(json.dumps(out))


def cpu_baseline(args, J, W0):
    """The CPU oracle (clean-room FP64 OpenMP restatement of rbm.cc's SR step,
    matrix-free S*v with O recomputed per MINRES iteration) on a bounded sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    threads = pyoracle.max_threads()
    C, m, B = args.cpu_C, args.cpu_m, args.B
    W = W0.copy()
    t0 = time.perf_counter()
    W, emat, info = pyoracle.sr_step(args.N, args.M, args.K, args.h, J, W, args.seed, C, m, B, args.lr,
                                     args.reglam, args.rtol)
    dt = time.perf_counter() - t0
    return {
        "value": C * m / dt,
        "unit": "MC samples/s (whole SR step)",
        "cores": threads,
        "kind": "port",
        "sample": f"one SR step with C={C} chains x m={m} samples, B={B}, same N/M/K/h/J/weights; "
                  f"{dt:.1f}s, {info['iters']} MINRES-QLP iterations",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--K", type=int, default=2)
    ap.add_argument("--C", type=int, default=4096)
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--B", type=int, default=64)
    ap.add_argument("--h", type=float, default=0.595)
    ap.add_argument("--lr", type=float, default=0.1)
    ap.add_argument("--reglam", type=float, default=40.0)
    ap.add_argument("--rtol", type=float, default=1e-9)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-C", type=int, default=512)
    ap.add_argument("--cpu-m", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (HBM counter calibration)")
    ap.add_argument("--workload", default="srstep", choices=["srstep", "scan", "serving"])
    ap.add_argument("--slots", type=int, default=8, help="scan/serving: concurrent solves")
    ap.add_argument("--solve-steps", type=int, default=5, help="scan/serving: SR steps per solve at most")
    ap.add_argument("--requests", type=int, default=48, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=4.0, help="serving: mean arrivals per second")
    args = ap.parse_args()

    if args.copy_calibration:
        # a plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE of this dispatch
        # give the counter scale used to turn rocprofv3 readings into HBM bytes
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4}))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is a "
                                   "design note only)"}))
        sys.exit(2)

    if args.workload != "srstep":
        serve_workload(args)
        return

    import nqes
    prec = nqes.PREC_FP64 if args.precision == "fp64" else nqes.PREC_MIXED
    J = load_J(args.N)
    W0 = initial_weights(args.N, args.M, args.K, args.seed)
    solver = nqes.Solver(args.N, args.M, args.K, args.h, J, precision=prec)
    W = W0.copy()
    S = args.C * args.m
    reglam, lr = args.reglam, args.lr
    step_i = 0

    def one_step():
        nonlocal reglam, lr, step_i
        emat, st = solver.monte_carlo(W, args.seed + step_i * args.C, args.C, args.m, args.B, max(lr, 0.01),
                                      max(reglam, 1.0), args.rtol, stats=True)
        reglam *= 0.9
        lr *= 0.99
        step_i += 1
        return emat, st

    for _ in range(args.warmup):
        one_step()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        emat, st = one_step()
        stats.append(st)
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps

    P = args.K * (args.N + args.M + args.N * args.M)
    real = not np.any(W0.imag)
    sv_env = os.environ.get("NQES_SV")
    factored = args.precision == "mixed" and args.N <= 64 and sv_env != "stored"
    q8 = factored and real and sv_env != "fp64mfma"
    if args.precision == "fp64":
        # two streamed passes over complex128 O, w written and read, v / result / o
        b_O, n_pass = 16, 2
        sv_bytes = n_pass * S * P * b_O + 2 * S * 16 + 3 * P * 16
        sv_kernel = "S*v: k_sv_rowdot64 + k_sv_colacc64 (two passes over complex128 O)"
    elif q8:
        # int8 form: each pass reads the quantised factors once (3 digit bytes per row
        # and column), plus per row its scale (8 B) and spin words, and w once per sample
        R = S * args.K
        ucols = 16 * ((args.K * args.M + args.K + 15) // 16)
        nw = (args.N + 31) // 32
        sv_bytes = 2 * R * ucols * 3 + 2 * R * 8 + R * 4 * nw + 2 * S * 8
        nsp = 32 if args.N <= 32 else 64
        nst = (args.N + 15) // 16
        ncg64 = (ucols + 63) // 64
        gemm = args.N <= 32 and ncg64 <= 4 and os.environ.get("NQES_Q8_FWD") != "tiles"
        # int8 MFMA ops issued: backward 18 16x16x64 MFMAs per 64 rows, column tile and
        # spin tile; forward either 6 per 16 rows and column tile (tiled form) or, as one
        # GEMM, 3 (6 nst + 1) 16x16x64 MFMAs per 16 rows and 64 columns
        bwd_ops = 2 * R * ucols * 18 * 16 * nst
        fwd_ops = R * ncg64 * 3 * (6 * nst + 1) * 2 * 16 * 64 if gemm else 2 * R * ucols * 6 * nsp
        sv_ops = fwd_ops + bwd_ops
        fwd_name = "k_q8_fwdg" if gemm else "k_q8_fwd"
        sv_kernel = f"S*v: {fwd_name} + k_q8_bwd (int8 MFMA over quantised c*tanh(theta) factors)"
    elif factored:
        # two FP64-MFMA passes over the factors: sigma^T B (R x N x (K M + K)) and the
        # transposed product, R = S K rows; complex weights double the real products
        R = S * args.K
        sv_flops = 2 * 2 * R * args.N * (args.K * args.M + args.K) * (1 if real else 2)
        sv_kernel = "S*v: k_svf_fwd + k_svf_bwd (FP64 MFMA over spins and c*tanh(theta) factors)"
    else:
        # one streamed pass over FP32 (real) or complex64 O, v / result / o in FP64
        b_O, n_pass = (4 if real else 8), 1
        sv_bytes = n_pass * S * P * b_O + 3 * P * (8 if real else 16)
        sv_kernel = f"S*v: k_sv_rows_mx (one pass over {'FP32' if real else 'complex64'} O)"
    sv_calls = sum(s["sv_calls"] for s in stats)
    sv_ms = sum(s["t_sv_ms"] for s in stats)
    sv_avg_ms = sv_ms / max(sv_calls, 1)
    if q8:
        achieved = sv_bytes / (sv_avg_ms * 1e-3) / 1e9 if sv_calls else 0.0
        roofline = {"bound": "hbm", "kernel": sv_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(args, real, "q8"),
                    "sv_avg_ms": sv_avg_ms, "sv_bytes": sv_bytes,
                    "int8_mfma_tops": sv_ops / (sv_avg_ms * 1e-3) / 1e12 if sv_calls else 0.0,
                    "int8_mfma_peak_tops": INT8_MFMA_PEAK_TOPS}
    elif factored:
        achieved = sv_flops / (sv_avg_ms * 1e-3) / 1e12 if sv_calls else 0.0
        roofline = {"bound": "mfma", "kernel": sv_kernel, "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK_TFLOPS,
                    "traffic": profiled_traffic(args, real, "factored"), "sv_avg_ms": sv_avg_ms, "sv_flops": sv_flops}
    else:
        achieved = sv_bytes / (sv_avg_ms * 1e-3) / 1e9 if sv_calls else 0.0
        roofline = {"bound": "hbm", "kernel": sv_kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(args, real, "stored"),
                    "sv_avg_ms": sv_avg_ms, "sv_bytes": sv_bytes}
    energies = np.sort(np.linalg.eigvals(emat / args.N).real).tolist()
    out = {
        "metric": "MC samples/s (whole SR step); converged solves/hour; p50/p99 solve latency",
        "value": S * args.steps / dt,
        "unit": "MC samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (random N(0,0.01^2) real weights, seeded)",
        "config": {"workload": f"trapped-ion SR step N={args.N} M={args.M} K={args.K} h={args.h} "
                               f"C={args.C} m={args.m} B={args.B} rtol={args.rtol} ({args.precision})",
                   "N": args.N, "M": args.M, "K": args.K, "chains": args.C, "samples_per_chain": args.m,
                   "burn_in": args.B, "precision": args.precision},
        "solves_per_hour_200_steps": 3600e3 / (200 * ms_per_step),
        "breakdown_ms_per_step": {
            "sample": sum(s["t_sample_ms"] for s in stats) / args.steps,
            "moments": sum(s["t_moments_ms"] for s in stats) / args.steps,
            "solve": sum(s["t_solve_ms"] for s in stats) / args.steps,
            "sv_device": sv_ms / args.steps,
        },
        "minres_iters_per_step": [s["minres_iters"] for s in stats],
        "accept_rate": sum(s["accepts"] for s in stats) / max(sum(s["proposals"] for s in stats), 1),
        "energies_last_step": energies,
        "roofline": roofline,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, J, W0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
