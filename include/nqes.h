/*
 * nqes.h -- C ABI of the MI355X multi-RBM variational Monte Carlo solver
 * (K lowest states of the long-range trapped-ion Ising model).
 *
 * The library (libnqes.so) replaces the in-process hot path of the reference
 * executable rbm-matfree-N*-M*-K* (reference file code/rbm.cc): one
 * stochastic-reconfiguration (SR) step = Metropolis sampling, local energies,
 * log-derivatives O, the matrix-free S*v operator, MINRES-QLP and the weight
 * update.  Everything is runtime-sized (the reference compiles N, M, K and J in).
 *
 * Conventions
 *   complex numbers are interleaved doubles (re, im), i.e. C99 double _Complex;
 *   weights W hold P = K*L complex numbers, L = N + M + N*M; RBM k occupies
 *     [k*L, (k+1)*L): a[N], b[M], w[N][M] (i-major) -- the index layout of
 *     the reference's O vector (rbm.cc:216, 223, 234);
 *   J is N x N row-major FP64 (symmetric, zero diagonal);
 *   spins: NW = ceil(N/32) 32-bit words per replica, bit i = spin i, a set bit
 *     means sigma_i = -1 (rbm.cc:141); per sample K*NW words, replica-major;
 *   K x K matrices (e_mat, H_loc) are row-major [replica k][RBM kk].
 * Errors: every call returns an int status (NQES_OK = 0); nqes_last_error()
 * gives a message.  No C++ exception crosses the ABI.  Calls on one context
 * must be serialised; use one context per host thread.  Each context owns one
 * HIP stream on its device.
 */
#ifndef NQES_H
#define NQES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NQES_OK = 0,
    NQES_EINVAL = 1,     /* invalid argument or unsupported size            */
    NQES_ENOMEM = 2,     /* device allocation failed / does not fit in HBM  */
    NQES_EDEVICE = 3,    /* HIP runtime error                               */
    NQES_ENONFINITE = 4, /* non-finite energy or gradient (rbm.cc's SIGFPE) */
    NQES_ESTATE = 5,     /* call order violated (e.g. solve before sample)  */
    NQES_ECANCELED = 6   /* solve cancelled through nqes_server_cancel       */
};

enum {
    NQES_PREC_FP64 = 0,  /* FP64 everywhere; bit-exact with the CPU oracle's
                            canonical order for spins, H_loc and O          */
    NQES_PREC_MIXED = 1  /* FP32 amplitudes/ratios/O, FP64 accumulation,
                            FP64 S*v reductions and MINRES-QLP              */
};

enum {
    NQES_MODEL_TRAPPEDION = 0, /* ising-like (rbm.cc:1042-1064): E_zz = 2 sum_{i>j} J_ij s_i s_j
                                  (rbm.cc:1012-1031) and the field -h sum_i sigma^x_i; also
                                  the ising, periodicising and longrange models with their J */
    NQES_MODEL_HS = 1,         /* hs-like (rbm.cc:1092-1119) with delta = 1:
                                  sum_{i>j} J_ij (XX + YY + delta ZZ), h unused; J = hs_model
                                  (rbm.cc:1066-1078) for the Haldane-Shastry chain */
    NQES_MODEL_XXZ = 2         /* hs-like with delta = -1; J = xxz_model (rbm.cc:1083-1089) */
};

typedef struct nqes_ctx nqes_ctx;

typedef struct {
    int minres_flag;          /* MINRES-QLP flag (minresqlp.h:14-54)          */
    uint64_t minres_iters;    /* Lanczos iterations                           */
    uint64_t accepts;         /* accepted Metropolis proposals, all chains    */
    uint64_t proposals;       /* C * (B + m) * K * N                          */
    uint64_t sv_calls;        /* S*v operator applications                    */
    double t_sample_ms;       /* sampling + H_loc + O kernel                  */
    double t_moments_ms;      /* column reductions o, |O|^2, E O*             */
    double t_solve_ms;        /* MINRES-QLP including all S*v                 */
    double t_sv_ms;           /* device time of all S*v kernels               */
    double t_total_ms;        /* whole SR step (host wall clock)              */
    int nonfinite;            /* 1 if a non-finite energy/gradient was seen   */
    uint64_t ulp_ties;        /* FP64 path: accept decisions whose uniform draw u
                                 lies within 2^-40 min(1, p) of min(1, p), p the
                                 amplitude ratio |det Phi'|^2 / |det Phi|^2
                                 (rbm.cc:540-543): decisions that an implementation
                                 rounding p in another order may take the other
                                 way; 0 on the mixed path (not counted)         */
} nqes_stats;

/* Creates a context on HIP device `device`.  Returns NULL on failure. */
nqes_ctx *nqes_create(int device);
void nqes_destroy(nqes_ctx *ctx);
/* ctx == NULL: the calling thread's last nqes_apply_S_rows_noalloc error. */
const char *nqes_last_error(const nqes_ctx *ctx);

/*
 * Sets the problem: N spins, M hidden units (M % N == 0 as rbm.cc:148
 * requires), K states, field h, couplings J (N x N), model, precision.
 * Replaces the compile-time NUM_SPINS/NUM_HIDDEN_UNITS/NUM_EIGENSTATES/JMAT_INC
 * of the reference (code/build.sh:28-42, rbm.cc:1015).
 */
int nqes_problem_init(nqes_ctx *ctx, int N, int M, int K, double h, const double *J,
                      int model, int precision);

/*
 * One SR step: monte_carlo<N,M,K>(rbm, seed, C, m, B, lr, reglam, rtol, maxit, H)
 * (rbm.cc:564-720).  Chain c uses pcg64(seed + c).  W (P complex) is updated in
 * place (w -= lr * x); emat (K*K complex) receives the sample mean of H_loc
 * (before the division by N that run() applies, rbm.cc:1183).  stats may be NULL.
 */
int nqes_sr_step(nqes_ctx *ctx, double *W, uint64_t seed, int C, int m, int B,
                 double lr, double reglam, double rtol, uint64_t maxit,
                 double *emat, nqes_stats *stats);

/*
 * The SR step in parts (used by parity tests and the benchmark).
 * nqes_sample: sampling, local energies, O written to device memory, and the
 *   normalised moments o, skk = reglam (<|O|^2> - |o|^2), f = <E O*> - <E><O*>
 *   (rbm.cc:596-650).  Keeps them in the context.
 * nqes_get_samples: per-sample spins (S*K*NW words), H_loc (S*K*K complex) and
 *   per-chain accepted-proposal counts (C).  Any pointer may be NULL.
 * nqes_get_O: rows [row0, row0+nrows) of O (row = chain*m + sample), P complex each.
 * nqes_get_moments: o, skk, f (P complex each), eloc (1 complex), emat (K*K complex).
 * nqes_apply_S: out = S v as the reference's matrix-free lambda (rbm.cc:661-684).
 * nqes_solve: MINRES-QLP (minresqlp.h:114-456) on S x = f; x: P complex.
 */
int nqes_sample(nqes_ctx *ctx, const double *W, uint64_t seed, int C, int m, int B,
                double reglam);
int nqes_get_samples(nqes_ctx *ctx, uint32_t *spins, double *H, uint64_t *accepts);
int nqes_get_O(nqes_ctx *ctx, int64_t row0, int64_t nrows, double *out);
int nqes_get_moments(nqes_ctx *ctx, double *o, double *skk, double *f, double *eloc,
                     double *emat);
/* Near-tie accept decisions of the last nqes_sample (FP64 path; see nqes_stats.ulp_ties):
 * per chain (C entries, may be NULL) and their total (may be NULL).  The reference
 * counterpart is the accept test of monte_carlo_step (rbm.cc:540-543), which has no
 * such count; 0 on the mixed path. */
int nqes_get_ulp_ties(nqes_ctx *ctx, uint64_t *per_chain, uint64_t *total);
/* Launch shape of the last mixed-precision nqes_sample (the reference runs one chain per
 * OpenMP thread, rbm.cc:564-600, and has no counterpart): Metropolis chains per block of
 * the sweep kernel, whether its tables were staged in LDS (1) or read from L2 (0), and
 * samples per block of the measurement kernel.  These shapes are chosen by the library's
 * tuner: after one untimed call with the default shapes, the next sampling calls of each
 * problem shape (device, real or complex weights, N, M, K, C, m, B, model) in the process
 * try candidate shapes in turn, timed on the device, and later calls use the fastest.  No lane computes anything else under another
 * shape, so results do not depend on the choice.  tuning: 1 while the shape is being tuned,
 * 0 once its choice is fixed, -1 with the tuner off (NQES_TUNE=0, or NQES_WPB / NQES_TABLDS
 * set).  Any pointer may be NULL. */
int nqes_get_sampler_launch(nqes_ctx *ctx, int *chains_per_block, int *tables_in_lds,
                            int *samples_per_block, int *tuning);
int nqes_apply_S(nqes_ctx *ctx, const double *v, double *out);

/*
 * No counterpart in the reference: the variant of nqes_apply_S for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The FP64 S*v of the matrix-free operator (rbm.cc:661-684)
 *   out = (1/S) sum_s conj(O_s) (O_s^T v) + skk .* v - (o^T v) conj(o)
 * over O rows the caller holds in device memory, by the kernels nqes_apply_S runs on
 * the FP64 path.  O: S rows, row s at O + s*ldo entries; cplx = 0: real O (doubles,
 * ldo even), cplx = 1: complex O (interleaved pairs); ldo >= P, entries past P finite.
 * o, skk, v, out: P entries of the same kind, device memory of the current device.
 * Enqueued on `stream` (a hipStream_t; NULL: the default stream) without waiting.
 * Allocates nothing: its partial rows and sums live in one static device array of
 * 24 MiB per device (each device holds it until the process ends); shapes whose rows
 * exceed the one-pass kernel (v in LDS, at most 16 x 64 x 9 16-byte vectors) or whose
 * scratch would exceed the array return NQES_EINVAL, with the reason in
 * nqes_last_error(NULL).  Calls that overlap on one device share the array and race.
 */
int nqes_apply_S_rows_noalloc(const double *O, int64_t S, int64_t P, int64_t ldo, int cplx,
                              const double *o, const double *skk, const double *v, double *out,
                              void *stream);
int nqes_solve(nqes_ctx *ctx, double rtol, uint64_t maxit, double *x, int *flag,
               uint64_t *iters);

/*
 * Correlation estimators of the final state: correlation_fcn (rbm.cc:730-771) with
 * the operators corrfcn_x/z/xx/zz (rbm.cc:774-829).  Chains as in an SR step
 * (pcg64(seed + c), B burn-in sweeps, m samples); per sample the K x K matrix
 * Phi^+ fcn(idx) is averaged.  Outputs (K*K complex each, row-major): x[N], z[N]
 * (index i), xx[N*N], zz[N*N] (index i*N + j).  The caller projects them on the
 * eigenvectors of e_mat as rbm.cc:764-769 does.  Runs the FP64 sampler.
 */
int nqes_correlations(nqes_ctx *ctx, const double *W, uint64_t seed, int C, int m, int B,
                      double *x, double *z, double *xx, double *zz);

/* Device bytes nqes_sample needs for C chains x m samples (admission check), an upper
 * bound that assumes complex weights.  For NQES_PREC_MIXED it counts the factored
 * S*v state (spins and c*tanh(theta) rows, no O) when N <= 64 and NQES_SV is not
 * "stored", else the stored O rows. */
int64_t nqes_required_bytes(int N, int M, int K, int C, int m, int precision);

/*
 * S form of the following SR steps: 0 = matrix-free (default; rbm-matfree-N*,
 * rbm.cc:661-684), 1 = dense P x P matrix oo = <conj(O) O^T> - conj(o) o^T formed
 * after sampling and applied as skk .* v + oo v (the rbm-N* executables,
 * rbm.cc:651-657, 686-696).  Dense needs NQES_PREC_FP64 and P*P*16 device bytes.
 */
int nqes_set_dense_S(nqes_ctx *ctx, int dense);

/*
 * Exhaustive evaluation behind the reference's --dump-states/ham/deriv files
 * (dump_states, rbm.cc:474-526, 1149-1172; needs N*K < 32): for every multi-spin
 * configuration cfg in [cfg0, cfg0 + ncfg) (ncfg <= 2^24; replica k's spins are bits
 * [k N, k N + N) of cfg) with weights W, on the FP64 path: phi (ncfg x K*K complex,
 * state_fcn), res (ncfg x K*K: the local-energy numerator of the model, H_loc =
 * Phi^+ res, rbm.cc:1042-1064 / 1092-1119) and tanh (ncfg x K*K*M, tanh(theta) of RBM
 * kk at replica k, [k][kk][j]).  Any output may be NULL.  Invalidates the samples.
 */
int nqes_enumerate(nqes_ctx *ctx, const double *W, int64_t cfg0, int64_t ncfg, double *phi, double *res,
                   double *tanh_);

/*
 * Solve server: continuous batching of whole solves on one device.
 *
 * A solve is the reference's run() (rbm.cc:1173-1197): up to niters SR steps with
 * the lrate/reglam schedules, chain seeds seed + t*C + c for step t, and its
 * convergence test; each request carries its own N, M, K, h, J, model and
 * precision (the reference compiles one binary per N, M, K, J).  Requests queue in
 * arrival order and are admitted while fewer than max_slots solves run and the
 * device bytes they need (nqes_required_bytes plus a fixed per-solve allowance)
 * fit in the budget.  A request that does not fit waits; later requests that fit
 * may pass it, at most 2*max_slots times, after which admission waits for it.
 * Every admitted solve runs on its own HIP stream, so the kernels of concurrent
 * solves overlap on the device.  A solve's result depends only on its request:
 * the same request gives the same bits alone and in any mix.
 */
typedef struct nqes_server nqes_server;

typedef struct {
    int N, M, K;              /* as nqes_problem_init                          */
    double h;
    const double *J;          /* N x N row-major; copied at submit             */
    int model, precision;
    const double *W0;         /* initial weights, P complex; copied at submit  */
    uint64_t seed;            /* chain c of SR step t: pcg64(seed + t*C + c)   */
    int C, m, B;              /* chains, samples per chain, burn-in sweeps     */
    int niters;               /* maximum SR steps (-n)                         */
    double tol;               /* convergence tolerance (--tol)                 */
    double lrate_init, lrate_decay, lrate_cutoff;
    double reglam_init, reglam_decay, reglam_cutoff;
    double rtol;              /* MINRES-QLP relative tolerance                 */
    uint64_t maxit;           /* MINRES-QLP iteration limit                    */
} nqes_request;

typedef struct {
    int status;               /* NQES_OK or the status of the failed call      */
    int steps;                /* SR steps taken                                */
    int converged;            /* 1 if the convergence test ended the solve     */
    int K;
    double energies[4];       /* ascending E/N of the last step (K entries)    */
    double t_submit_ms;       /* server clock: ms since nqes_server_create     */
    double t_admit_ms;
    double t_done_ms;
    double t_device_ms;       /* sum of the SR steps' wall times               */
    int64_t bytes;            /* device bytes reserved for the solve           */
} nqes_result;

/* budget <= 0: 90 % of the device memory free at creation.  max_slots >= 1.
 * Returns NULL on failure (no device, bad arguments). */
nqes_server *nqes_server_create(int device, int max_slots, int64_t budget);
/* Waits for every queued and running solve, then frees the server. */
void nqes_server_destroy(nqes_server *srv);
const char *nqes_server_last_error(const nqes_server *srv);
/* Validates and queues a request; *id names it.  NQES_EINVAL for bad sizes,
 * NQES_ENOMEM if it can never fit in the budget. */
int nqes_server_submit(nqes_server *srv, const nqes_request *req, uint64_t *id);
/* *done = 1 once solve `id` has finished. */
int nqes_server_poll(nqes_server *srv, uint64_t id, int *done);
/* *steps = SR steps solve `id` has completed so far (0 while it is queued), for
 * progress reports; the reference prints one energy line per SR step instead
 * (rbm.cc:1188-1190). */
int nqes_server_progress(nqes_server *srv, uint64_t id, int *steps);
/* Blocks until solve `id` finishes; copies its final weights (P complex, may be
 * NULL) and result, then forgets the solve.  Returns the solve's status. */
int nqes_server_wait(nqes_server *srv, uint64_t id, double *W, nqes_result *res);
/* Cancels solve `id` (no counterpart in the reference, whose run() is one solve per
 * process: a serving deadline).  A queued request finishes at once, a running one
 * after its current SR step; either way with status NQES_ECANCELED and the weights and
 * energies of its last completed step.  A finished solve is left as it is.  The id
 * stays valid for nqes_server_wait. */
int nqes_server_cancel(nqes_server *srv, uint64_t id);

/* Library version string. */
const char *nqes_version(void);

#ifdef __cplusplus
}
#endif

#endif /* NQES_H */
