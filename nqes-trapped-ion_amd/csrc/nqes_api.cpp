// nqes_api.cpp -- host side of libnqes: the C ABI of include/nqes.h.
//
// One SR step (reference: monte_carlo, code/rbm.cc:564-720):
//   1. weights -> device, per-step tables
//   2. sampler kernel: Metropolis chains, H_loc, E_loc and the O matrix (S x P) in HBM
//   3. deterministic column reductions -> sums of O, |O|^2, E O*, H_loc, E_loc
//   4. host normalisation (rbm.cc:641-650)
//   5. MINRES-QLP on the host (FP64 scalars, length-P vectors) whose operator is the
//      streamed S*v over O on the device (rbm.cc:661-684)
//   6. w -= lr * x (rbm.cc:699-717)
// Precision NQES_PREC_FP64 follows the canonical FP64 order (bit-exact spins, H_loc,
// O); NQES_PREC_MIXED keeps per-hidden-unit work and O in FP32 and all sums in FP64,
// and specialises to real arithmetic when every weight is real.
// Host arithmetic is compiled with -ffp-contract=off as the FP64 device code is.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nqes.h"
#include "cplx64.hpp"
#include "kernels.hpp"
#include "minresqlp.hpp"

using nq64::cd;

std::shared_mutex &nq_capture_mutex()
{
    static std::shared_mutex m;
    return m;
}

bool &nq_capturing()
{
    thread_local bool capturing = false;
    return capturing;
}

namespace {

// a synchronous copy, excluded from stream captures on other threads
hipError_t copy_sync(void *dst, const void *src, size_t n, hipMemcpyKind kind)
{
    NqUnsafeHold hold;
    return hipMemcpy(dst, src, n, kind);
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            NqUnsafeHold hold;
            (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) n = 16;
        NqUnsafeHold hold;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return e;
        }
        bytes = n;
        return hipSuccess;
    }
};

int choose_chunks(long long S, long long ncol, int cols_per_block)
{
    long long colblocks = (ncol + cols_per_block - 1) / cols_per_block;
    long long nch = 4096 / (colblocks > 0 ? colblocks : 1);
    if (nch < 1) nch = 1;
    if (nch > S) nch = S > 0 ? S : 1;
    if (nch > 4096) nch = 4096;
    return (int)nch;
}

// ---- launch-shape tuner of the mixed sampler.  The sweeps' chains per block (wavefronts
// per block), the placement of their tables (LDS or L2) and the measurement kernel's
// samples per block spread the same per-lane work over the CUs differently: no lane
// computes anything else, so spins, local energies and factor rows are the same bit for
// bit for every choice.  The first sampling calls of each problem shape in the process
// (device, real or complex weights, N, M, K, C, m, B, model) try the candidates in turn,
// timed by the sampling events, after one untimed call with the default shapes (a first
// call pays one-time costs: code loading, cold caches); later calls take the fastest
// (ties: the earlier candidate).  NQES_TUNE=0, NQES_WPB or NQES_TABLDS turn it off.
struct SamplerChoice {
    int sweep_wpb, sweep_tab, measure_wpb;
};
struct TuneEntry {
    std::vector<SamplerChoice> cand;
    std::vector<double> ms;  // < 0: not timed yet
    bool warm = false;       // the shape's first call (default shapes) has run, untimed
    size_t next = 0;         // next candidate to hand out
    int best = -1;           // fixed once every candidate is timed
};
using TuneKey = std::array<long long, 9>;
std::mutex g_tune_mu;
std::map<TuneKey, TuneEntry> g_tune;

bool tuner_on()
{
    const char *t = getenv("NQES_TUNE"), *w = getenv("NQES_WPB"), *l = getenv("NQES_TABLDS");
    return !(t && *t && atoi(t) == 0) && !(w && *w) && !(l && *l);
}

// the candidates around the default shapes: sweep wavefronts per block wpb, wpb / 2; the
// tables from L2 as well when the default puts them in LDS; measurement 4 or 2 wavefronts
std::vector<SamplerChoice> tune_candidates(const SweepShape &sw, const MeasureShape &ms)
{
    std::vector<int> swpb{sw.wpb}, tabs{sw.tab}, mwpb{ms.wpb};
    if (sw.wpb > 1) swpb.push_back(sw.wpb / 2);
    if (sw.tab) tabs.push_back(0);
    if (ms.wpb > 1) mwpb.push_back(ms.wpb / 2);
    std::vector<SamplerChoice> out;  // the default shapes first
    for (int w : swpb)
        for (int t : tabs)
            for (int mw : mwpb) out.push_back(SamplerChoice{w, t, mw});
    return out;
}

int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return v && *v ? atoi(v) : dflt;
}

} // namespace

struct nqes_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    size_t lds_max = 65536;

    // problem
    bool has_problem = false;
    int N = 0, M = 0, K = 0, NW = 0, L = 0, model = 0, prec = 0;
    long long P = 0;
    double h = 0;
    DevBuf dJ, dW, dchT, dshT, deaP, deaN;
    DevBuf dwT, dbT, daD, dspins0, drng0, dtheta0, dflags;  // mixed path
    DevBuf dtabp, dtabm;                               // its sampler's padded tables, the measurement's copy

    // sample state
    bool sampled = false;
    bool cplx = true;          // mixed path: complex arithmetic (false: all weights real)
    int C = 0, m = 0, B = 0;
    long long S = 0, Pld = 0;
    DevBuf dO, dHs, dEs, dspins, dacc, dties;
    DevBuf dmpart_o, dmpart_s, dmpart_f, dmo, dms, dmf, dhpart, dhsum, depart, desum;
    DevBuf dv, dov, dw, dsvpart, dsvwpart, dsvout, dsvscratch, dmscratch;
    DevBuf dU, dwrow, dsvfpart, dsvfscratch, dsvftmp;  // factored S*v and moments (mixed path)
    DevBuf dUq, descale, dvq, dq8w, dq8m;              // its int8 form (real weights)
    DevBuf dUf;                                        // row-major digits (GEMM forward pass)
    DevBuf dw3, dvin3, dsvout3;                        // complex weights, int8 form: per-set sample sums,
                                                       // operands in, the three raw sums out
    DevBuf dUmax;                                      // max |U| per factor row
    DevBuf dxc;                                        // int8 form: the c columns' integers (kept apart)
    hipGraph_t svg = nullptr;                          // int8 S*v sequence, captured once per
    hipGraphExec_t svx = nullptr;                      // sample set and replayed per MINRES step
    double *hv_pin = nullptr, *hr_pin = nullptr;       // pinned v | o^T v and result (P + 1 each)
    size_t pin_len = 0;
    bool dense = false;                                // dense S (FP64): oo P x P
    DevBuf doo, dvo, dskk;
    DevBuf denum;                                      // exhaustive evaluation outputs
    bool svf = false, q8 = false, q8c = false;        // q8c: complex weights on the int8 form
    bool oreal = false;                                // FP64 path, real weights: O stored as doubles
    Sv64Shape sv64{};                                  // FP64 path: one-pass S*v shape (ok = 0: two passes)
    DevBuf dsvwsum;
    SvfShape svf_shape{};
    Q8Shape q8_shape{};
    int mchunks = 1, svchunks = 1;
    std::vector<cd> o, skk, f, emat;
    cd eloc{0, 0};
    std::vector<cd> svraw;
    std::vector<double> hv;
    uint64_t sv_calls = 0;
    double sv_ms = 0;
    int nonfinite = 0;
    int theta_overflow = 0;  // mixed path: an angle |theta| > 80 (FP32 cosh would overflow)
    // sampler launch of the last mixed nqes_sample and its tuner state (tune_idx >= 0: a
    // trial whose time nqes_sample records under tune_key)
    SamplerChoice launch{0, 0, 0};
    int tune_state = -1;  // 1: trial, 0: tuned choice, -1: tuner off
    int tune_idx = -1;
    TuneKey tune_key{};

    int fail(int code, const char *fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    int hipfail(hipError_t e, const char *what)
    {
        return fail(e == hipErrorOutOfMemory ? NQES_ENOMEM : NQES_EDEVICE, "%s: %s", what, hipGetErrorString(e));
    }
};

#define HIPCHK(ctx, expr, what)                                                                                  \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return (ctx)->hipfail(e_, what);                                                   \
    } while (0)

static double event_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}

// bytes of one O element and the padded row length (rows start on 16 bytes)
static size_t o_elem(int prec, bool cplx) { return prec == NQES_PREC_FP64 ? (cplx ? 16 : 8) : (cplx ? 8 : 4); }
static long long o_ld(long long P, int prec, bool cplx)
{
    size_t e = o_elem(prec, cplx);
    long long per16 = 16 / (long long)e;
    return (P + per16 - 1) / per16 * per16;
}

// the captured S*v graph refers to the sample set's buffers
static void drop_sv_graph(nqes_ctx *c)
{
    NqUnsafeHold hold;
    if (c->svx) (void)hipGraphExecDestroy(c->svx);
    if (c->svg) (void)hipGraphDestroy(c->svg);
    c->svx = nullptr;
    c->svg = nullptr;
}

// lane-group size, waves per block and table staging of a mixed kernel with G lanes
// per chain (sweeps) or per sample (measurement); the LDS layout is the same
extern "C" {

const char *nqes_version(void) { return "nqes-mi355x 0.2 (gfx950)"; }

nqes_ctx *nqes_create(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    nqes_ctx *c = new nqes_ctx();
    c->device = device;
    NqUnsafeHold hold;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return nullptr;
        }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->lds_max = prop.sharedMemPerBlock;
    (void)nq_noalloc_setup(device, nullptr, nullptr);  // the no-allocation entry point's scratch address
    return c;
}

void nqes_destroy(nqes_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    NqUnsafeHold hold;
    if (c->svx) (void)hipGraphExecDestroy(c->svx);
    if (c->svg) (void)hipGraphDestroy(c->svg);
    if (c->hv_pin) (void)hipHostFree(c->hv_pin);
    if (c->hr_pin) (void)hipHostFree(c->hr_pin);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nqes_last_error(const nqes_ctx *c) { return c ? c->err.c_str() : nq_noalloc_last_error(); }

int64_t nqes_required_bytes(int N, int M, int K, int C, int m, int precision)
{
    // the allocations of nqes_problem_init and nqes_sample, for complex weights
    // (real weights need less on the mixed path)
    if (N <= 0 || M <= 0 || K <= 0 || C <= 0 || m <= 0) return -1;
    if (precision != NQES_PREC_FP64 && precision != NQES_PREC_MIXED) return -1;
    const bool mixed = precision == NQES_PREC_MIXED, cplx = true;
    const long long P = (long long)K * (N + M + (long long)N * M);
    const long long S = (long long)C * m;
    const long long NW = (N + 31) / 32;
    const long long Pld = o_ld(P, precision, cplx);
    long long b = 8LL * N * N + 16 * P + 32LL * K * N * M + 32LL * K * N;  // J, W, tables
    if (mixed) b += 8LL * K * N * M + 8LL * K * M + 16LL * K * N + 16;
    if (mixed) b += (long long)C * K * NW * 4 + C * 16LL + (long long)C * K * K * M * 8;  // chain starts
    if (mixed) b += 2 * (long long)nq_sweep_shape(cplx, N, M, K, 0).tab_bytes;  // sampler tables, measurement copy
    SvfShape sh{};
    bool svf = false;
    if (mixed) {
        const char *e = getenv("NQES_SV");
        sh = nq_svf_shape(cplx, N, M, K);
        svf = sh.ok && !(e && strcmp(e, "stored") == 0);
    }
    const long long mch = choose_chunks(S, P, 256);
    if (svf) {  // factor rows, per-row sums and block partials instead of O
        const long long npair = K > 1 ? (long long)K * (K - 1) / 2 : 1;
        const long long nchunk = (long long)sh.bwd_blocks * npair;
        b += S * K * (long long)sh.ucols * 8;
        b += (long long)sh.fwd_cgs * S * K * 16;
        b += std::max((long long)sh.bwd_blocks * 16, nchunk * 8) * (P + 1);
        b += (nchunk / 32 + 2) * (P + 1) * 16 + 3 * (P + 1) * 16;
        // int8 form; complex weights keep three digit sets per row
        const Q8Shape q8 = nq_q8_shape(N, M, K, S * K);
        if (q8.ok) {
            b += 3 * (q8.rpad * (long long)q8.ucols * 3 + (long long)q8.uf_bytes + (long long)q8.xc_bytes) +
                 q8.rpad * 12 +
                 (long long)(q8.vq_bytes + q8.w_bytes + q8.m_bytes) + 3 * S * 8 + (3 * P + 4) * 8 + 3 * (P + 1) * 8;
            if (cplx) {  // the complex S*v's per-set buffers hold three sets
                size_t vq, q8w, part, scr;
                nq_q8c_sizes(nq_q8_shape(N, M, K, S * K, true), P, &vq, &q8w, &part, &scr);
                b += (long long)(vq + q8w + part + scr);
            }
        }
    } else {
        b += S * Pld * (long long)o_elem(precision, cplx);  // O
        if (mixed) b += (mch / 32 + 2) * P * 16;              // moments scratch
    }
    b += S * K * K * 16 + S * 16 + S * K * NW * 4 + C * 8LL;  // H_loc, E_loc, spins, accepts
    b += (svf ? 0 : 3 * mch * P * 16) + 3 * P * 16;           // moment partials and sums
    b += (long long)choose_chunks(S, (long long)K * K, 64) * K * K * 16 + K * K * 16;
    b += (long long)choose_chunks(S, 1, 64) * 16 + 16;
    b += P * 16 + 16 + S * 16 + (P + 1) * 16;  // v, o^T v, w, S*v result
    if (!mixed) {
        b += mch * P * 16;
        const Sv64Shape s64 = nq_sv64_shape(P, S, 160 * 1024);  // complex rows: the larger partials
        if (s64.ok) b += (long long)s64.blocks * (P * 16 + 16) + (s64.blocks / 32 + 2) * (P + 1) * 16;
    } else if (!svf) {
        const SvShape sv = nq_sv_shape(cplx, Pld / 2);
        long long nblk = sv.qv ? sv.blocks : std::min<long long>(mch, S);
        b += nblk * Pld * 16 + nblk * 16 + (nblk / 32 + 2) * (Pld + 1) * 16;
    }
    return b;
}

int nqes_problem_init(nqes_ctx *c, int N, int M, int K, double h, const double *J, int model, int precision)
{
    if (!c) return NQES_EINVAL;
    if (hipSetDevice(c->device) != hipSuccess) return c->fail(NQES_EDEVICE, "hipSetDevice failed");
    if (N < 1 || N > 256) return c->fail(NQES_EINVAL, "N=%d outside [1, 256]", N);
    if (M < 1 || M > 512) return c->fail(NQES_EINVAL, "M=%d outside [1, 512]", M);
    if (K < 1 || K > 4) return c->fail(NQES_EINVAL, "K=%d outside [1, 4]", K);
    if (M % N != 0) return c->fail(NQES_EINVAL, "M=%d must be a multiple of N=%d (rbm.cc:148)", M, N);
    if (!J) return c->fail(NQES_EINVAL, "J is NULL");
    if (model != NQES_MODEL_TRAPPEDION && model != NQES_MODEL_HS && model != NQES_MODEL_XXZ)
        return c->fail(NQES_EINVAL, "unsupported model %d", model);
    if (precision != NQES_PREC_FP64 && precision != NQES_PREC_MIXED)
        return c->fail(NQES_EINVAL, "unsupported precision %d", precision);
    if (precision == NQES_PREC_FP64) {
        size_t lds = (size_t)16 * K * K * M + 4 * (size_t)K * ((N + 31) / 32);
        if (lds > c->lds_max)
            return c->fail(NQES_EINVAL, "K*K*M=%d hidden-unit angles (%zu B) exceed the %zu B of LDS per block",
                           K * K * M, lds, c->lds_max);
    } else {
        const SweepShape sw = nq_sweep_shape(true, N, M, K, c->lds_max);
        if (!sw.ok || !nq_measure_shape(true, N, K, sw, c->lds_max).ok)
            return c->fail(NQES_EINVAL, "K*K*M=%d hidden-unit angles exceed the LDS per block", K * K * M);
    }
    c->N = N;
    c->M = M;
    c->K = K;
    c->NW = (N + 31) / 32;
    c->L = N + M + N * M;
    c->P = (long long)K * c->L;
    c->h = h;
    c->model = model;
    c->prec = precision;
    HIPCHK(c, c->dJ.ensure(sizeof(double) * N * N), "alloc J");
    HIPCHK(c, copy_sync(c->dJ.p, J, sizeof(double) * N * N, hipMemcpyHostToDevice), "copy J");
    HIPCHK(c, c->dW.ensure(16 * c->P), "alloc W");
    HIPCHK(c, c->dchT.ensure((size_t)16 * K * N * M), "alloc tables");
    HIPCHK(c, c->dshT.ensure((size_t)16 * K * N * M), "alloc tables");
    HIPCHK(c, c->deaP.ensure((size_t)16 * K * N), "alloc tables");
    HIPCHK(c, c->deaN.ensure((size_t)16 * K * N), "alloc tables");
    if (precision == NQES_PREC_MIXED) {
        HIPCHK(c, c->dwT.ensure((size_t)8 * K * N * M), "alloc tables");
        HIPCHK(c, c->dbT.ensure((size_t)8 * K * M), "alloc tables");
        HIPCHK(c, c->daD.ensure((size_t)16 * K * N), "alloc tables");
        HIPCHK(c, c->dflags.ensure(16), "alloc flags");
    }
    c->has_problem = true;
    c->dense = false;
    c->sampled = false;
    return NQES_OK;
}

static int sample_fp64(nqes_ctx *c, uint64_t seed, int C, int m, int B)
{
    const long long S = (long long)C * m;
    HIPCHK(c, nq_launch_tables64(c->N, c->M, c->K, c->L, c->dW.p, c->dchT.p, c->dshT.p, c->deaP.p, c->deaN.p, c->stream),
           "tables kernel");
    Fp64SampleArgs a{};
    a.N = c->N;
    a.M = c->M;
    a.NW = c->NW;
    a.L = c->L;
    a.P = c->P;
    a.C = C;
    a.m = m;
    a.B = B;
    a.seed = seed;
    a.h = c->h;
    a.model = c->model;
    a.delta = c->model == NQES_MODEL_XXZ ? -1.0 : 1.0;
    a.J = static_cast<const double *>(c->dJ.p);
    a.W = c->dW.p;
    a.chT = c->dchT.p;
    a.shT = c->dshT.p;
    a.eaP = c->deaP.p;
    a.eaN = c->deaN.p;
    a.O = c->svf ? nullptr : c->dO.p;
    a.Oreal = c->oreal ? 1 : 0;
    a.Pld = c->Pld;
    a.Hs = c->dHs.p;
    a.Es = c->dEs.p;
    a.spins = static_cast<uint32_t *>(c->dspins.p);
    a.accepts = static_cast<unsigned long long *>(c->dacc.p);
    a.ties = static_cast<unsigned long long *>(c->dties.p);
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream), "event");
    HIPCHK(c, nq_launch_sample64(c->K, a, c->stream), "sample kernel");
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream), "event");
    if (c->oreal)
        HIPCHK(c, nq_launch_moments64r(c->dO.p, c->Pld, c->dEs.p, S, c->P, c->mchunks, c->dmpart_o.p, c->dmpart_s.p,
                                       c->dmpart_f.p, c->dmo.p, c->dms.p, c->dmf.p, c->stream),
               "moments kernel");
    else
        HIPCHK(c, nq_launch_moments64(c->dO.p, c->dEs.p, S, c->P, c->mchunks, c->dmpart_o.p, c->dmpart_s.p,
                                      c->dmpart_f.p, c->dmo.p, c->dms.p, c->dmf.p, c->stream),
               "moments kernel");
    return NQES_OK;
}

// arguments of the factored S*v / moments kernels for S samples
static SvfArgs svf_args(nqes_ctx *c, long long S)
{
    SvfArgs f{};
    f.N = c->N;
    f.M = c->M;
    f.K = c->K;
    f.NW = c->NW;
    f.Ucols = c->svf_shape.ucols;
    f.n4 = (c->N + 3) & ~3;
    f.S = S;
    f.P = c->P;
    f.L = c->L;
    f.spins = static_cast<const uint32_t *>(c->dspins.p);
    f.U = c->dU.p;
    f.v = c->dv.p;
    f.ov = c->dov.p;
    f.wrow = c->dwrow.p;
    f.w = c->dw.p;
    f.part = c->dsvfpart.p;
    f.E = c->dEs.p;
    f.wsel = 0;
    f.ncg = 1;
    f.Uq = c->dUq.p;
    f.Uf = c->dUf.p;
    f.Umax = (const float *)c->dUmax.p;
    f.escale = static_cast<const double *>(c->descale.p);
    f.vq = c->dvq.p;
    f.q8w = c->dq8w.p;
    f.q8m = c->dq8m.p;
    f.qcols = (c->q8 || c->q8c) ? c->q8_shape.qcols : f.Ucols;
    f.xc = static_cast<float *>(c->dxc.p);
    return f;
}

// the int8 form's arguments for complex weights (the complex factor rows in U, the three
// digit sets in slices of Uq / Uf)
static SvfArgs q8c_args(nqes_ctx *c)
{
    SvfArgs f = svf_args(c, c->S);
    f.Umax = (const float *)c->dUmax.p;
    f.Ucols = c->q8_shape.ucols;
    return f;
}

// the sampler launch of this call: the tuned choice of its shape, the next untimed
// candidate while the shape is being tuned, or the default shapes (tuner off)
static void tune_sampler(nqes_ctx *c, int C, int m, int B, SweepShape &sw, MeasureShape &ms)
{
    c->tune_idx = -1;
    c->tune_state = -1;
    if (tuner_on()) {
        const TuneKey key{c->device, c->cplx ? 1 : 0, c->N, c->M, c->K, C, m, B, c->model};
        std::lock_guard<std::mutex> lk(g_tune_mu);
        TuneEntry &e = g_tune[key];
        if (e.cand.empty()) {
            e.cand = tune_candidates(sw, ms);
            e.ms.assign(e.cand.size(), -1.0);
        }
        int idx = e.best;
        if (idx < 0 && !e.warm) {  // the untimed first call
            idx = 0;
            e.warm = true;
            c->tune_state = 1;
        } else if (idx < 0 && e.next < e.cand.size()) {  // a trial
            idx = (int)e.next++;
            c->tune_idx = idx;
            c->tune_key = key;
            c->tune_state = 1;
        } else if (idx < 0) {  // every candidate handed out, not all timed yet: the fastest so far
            idx = 0;
            for (size_t q = 1; q < e.cand.size(); q++)
                if (e.ms[q] >= 0 && (e.ms[idx] < 0 || e.ms[q] < e.ms[idx])) idx = (int)q;
            c->tune_state = 1;
        } else {
            c->tune_state = 0;
        }
        sw.wpb = e.cand[idx].sweep_wpb;
        sw.tab = e.cand[idx].sweep_tab;
        ms.wpb = e.cand[idx].measure_wpb;
    }
    c->launch = SamplerChoice{sw.wpb, sw.tab, ms.wpb};
}

// a trial's sampling time (chain start, sweeps, measurement); the choice is fixed once
// every candidate of the shape has one
static void tune_record(nqes_ctx *c, double ms_sample)
{
    if (c->tune_idx < 0) return;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneEntry &e = g_tune[c->tune_key];
    if ((size_t)c->tune_idx < e.ms.size()) e.ms[c->tune_idx] = ms_sample > 0 ? ms_sample : 1e30;
    c->tune_idx = -1;
    if (e.best >= 0 || e.next < e.cand.size()) return;
    int best = 0;
    for (size_t q = 0; q < e.ms.size(); q++) {
        if (e.ms[q] < 0) return;
        if (e.ms[q] < e.ms[best]) best = (int)q;
    }
    e.best = best;
}

static int sample_mixed(nqes_ctx *c, uint64_t seed, int C, int m, int B)
{
    const long long S = (long long)C * m;
    const bool cplx = c->cplx;
    HIPCHK(c, c->dspins0.ensure((size_t)C * c->K * c->NW * 4), "alloc spins0");
    HIPCHK(c, c->drng0.ensure((size_t)C * 16), "alloc rng0");
    HIPCHK(c, c->dtheta0.ensure((size_t)C * c->K * c->K * c->M * (cplx ? 8 : 4)), "alloc theta0");
    HIPCHK(c, hipMemsetAsync(c->dflags.p, 0, 16, c->stream), "clear flags");
    HIPCHK(c, nq_launch_tables_mx(cplx, c->N, c->M, c->K, c->L, c->dW.p, c->dchT.p, c->dshT.p, c->dwT.p, c->dbT.p,
                                  c->deaP.p, c->deaN.p, c->daD.p, c->stream),
           "tables kernel");
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream), "event");
    HIPCHK(c, nq_launch_chain_init(c->N, c->K, c->NW, C, seed, (uint32_t *)c->dspins0.p, (uint64_t *)c->drng0.p,
                                   c->stream),
           "chain init kernel");
    HIPCHK(c, nq_launch_theta_gemm(cplx, c->N, c->M, c->K, c->NW, C, (const uint32_t *)c->dspins0.p, c->dwT.p,
                                   c->dbT.p, c->dtheta0.p, c->stream),
           "theta GEMM kernel");
    SweepShape sw = nq_sweep_shape(cplx, c->N, c->M, c->K, c->lds_max);  // k_sample_mx
    if (!sw.ok) return c->fail(NQES_EINVAL, "M=%d too large for the sampler", c->M);
    HIPCHK(c, c->dtabp.ensure(sw.tab_bytes), "alloc sampler tables");
    HIPCHK(c, c->dtabm.ensure(sw.tab_bytes), "alloc measurement tables");
    MeasureShape ms = nq_measure_shape(cplx, c->N, c->K, sw, c->lds_max);  // k_measure_mx
    if (!ms.ok) return c->fail(NQES_EINVAL, "K*K*M=%d hidden-unit angles exceed the LDS per block", c->K * c->K * c->M);
    tune_sampler(c, C, m, B, sw, ms);
    MixedSampleArgs a{};
    a.N = c->N;
    a.M = c->M;
    a.NW = c->NW;
    a.L = c->L;
    a.P = c->P;
    a.Pld = c->Pld;
    a.C = C;
    a.m = m;
    a.B = B;
    a.h = c->h;
    a.model = c->model;
    a.delta = c->model == NQES_MODEL_XXZ ? -1.0 : 1.0;
    a.J = static_cast<const double *>(c->dJ.p);
    a.chT = c->dchT.p;
    a.shT = c->dshT.p;
    a.wT = c->dwT.p;
    a.bT = c->dbT.p;
    a.eaP = c->deaP.p;
    a.eaN = c->deaN.p;
    a.aD = c->daD.p;
    a.theta0 = c->dtheta0.p;
    a.spins0 = (const uint32_t *)c->dspins0.p;
    a.rng0 = (const uint64_t *)c->drng0.p;
    a.O = c->svf ? nullptr : c->dO.p;
    a.Hs = c->dHs.p;
    a.Es = c->dEs.p;
    a.spins = static_cast<uint32_t *>(c->dspins.p);
    a.accepts = static_cast<unsigned long long *>(c->dacc.p);
    a.flags = static_cast<unsigned int *>(c->dflags.p);
    a.U = c->svf ? c->dU.p : nullptr;
    a.Ucols = c->svf_shape.ucols;
    a.Umax = c->q8 ? (float *)c->dUmax.p : nullptr;
    HIPCHK(c, nq_launch_sweep_mx(cplx, c->K, a, sw, c->dtabp.p, c->dtabm.p, c->stream), "mixed sample kernel");
    HIPCHK(c, nq_launch_measure_mx(cplx, c->K, a, ms, c->dtabm.p, c->stream), "mixed measure kernel");
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream), "event");
    if (c->svf) {
        SvfArgs f = svf_args(c, S);
        if (c->q8) {
            HIPCHK(c, nq_q8_pack(f, c->q8_shape, c->stream), "factor quantisation kernel");
            HIPCHK(c, nq_q8_moments(f, c->q8_shape, c->svf_shape, c->dsvfpart.p, c->dsvfscratch.p, c->dsvftmp.p,
                                    c->dmo.p, c->dms.p, c->dmf.p, c->stream),
                   "int8 moments kernels");
        } else {
            HIPCHK(c, nq_launch_svf_moments(cplx, f, c->svf_shape, c->dsvfpart.p, c->dsvfscratch.p, c->dsvftmp.p,
                                            c->dmo.p, c->dms.p, c->dmf.p, c->stream),
                   "factored moments kernels");
            if (c->q8c) {
                c->S = S;  // q8c_args sizes its rows by the sample count
                HIPCHK(c, nq_q8c_pack(q8c_args(c), c->q8_shape, c->stream), "complex factor quantisation kernels");
            }
        }
        return NQES_OK;
    }
    HIPCHK(c, c->dmscratch.ensure((size_t)(c->mchunks / 32 + 2) * c->P * 16), "alloc moments scratch");
    HIPCHK(c, nq_launch_moments_mx(cplx, c->dO.p, c->dEs.p, S, c->P, c->Pld, c->mchunks, c->dmpart_o.p,
                                   c->dmpart_s.p, c->dmpart_f.p, c->dmscratch.p, c->dmo.p, c->dms.p, c->dmf.p,
                                   c->stream),
           "moments kernel");
    return NQES_OK;
}

int nqes_sample(nqes_ctx *c, const double *W, uint64_t seed, int C, int m, int B, double reglam)
{
    if (!c) return NQES_EINVAL;
    if (!c->has_problem) return c->fail(NQES_ESTATE, "nqes_problem_init not called");
    if (!W) return c->fail(NQES_EINVAL, "W is NULL");
    if (C < 1 || m < 1 || B < 0) return c->fail(NQES_EINVAL, "invalid C=%d m=%d B=%d", C, m, B);
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    c->sampled = false;
    const int K = c->K;
    const long long P = c->P, S = (long long)C * m;
    bool cplx = false;
    for (long long q = 0; q < P; q++)
        if (W[2 * q + 1] != 0.0) {
            cplx = true;
            break;
        }
    c->cplx = cplx;
    // FP64 path: real weights make O real (every imaginary part is an exact zero), so
    // it is stored as doubles when the one-pass S*v applies to its rows; the dense S
    // form keeps complex rows
    c->oreal = false;
    c->sv64 = Sv64Shape{0, 0, 0, 0};
    if (c->prec == NQES_PREC_FP64 && !c->dense) {
        if (!cplx) {
            c->sv64 = nq_sv64_shape(o_ld(P, NQES_PREC_FP64, false) / 2, S, c->lds_max);
            c->oreal = c->sv64.ok != 0;
        }
        if (!c->oreal) c->sv64 = nq_sv64_shape(P, S, c->lds_max);
    }
    c->Pld = o_ld(P, c->prec, c->prec == NQES_PREC_FP64 ? !c->oreal : cplx);
    drop_sv_graph(c);  // buffers may move
    // factored S*v (mixed path) unless NQES_SV=stored or the shape does not apply; for
    // real weights its int8 form unless NQES_SV=fp64mfma
    c->svf = c->q8 = c->q8c = false;
    if (c->prec == NQES_PREC_MIXED) {
        const char *e = getenv("NQES_SV");
        c->svf_shape = nq_svf_shape(cplx, c->N, c->M, K);
        c->svf = c->svf_shape.ok && !(e && strcmp(e, "stored") == 0);
        const bool fp64mfma = e && strcmp(e, "fp64mfma") == 0;
        if (c->svf && !cplx) {
            c->q8_shape = nq_q8_shape(c->N, c->M, K, S * K);
            c->q8 = c->q8_shape.ok && !fp64mfma;
        } else if (c->svf) {  // complex: the int8 form with three digit sets per row
            c->q8_shape = nq_q8_shape(c->N, c->M, K, S * K, true);
            c->q8c = c->q8_shape.ok && !fp64mfma;
        }
    }
    if (c->svf) {
        const size_t tsz = cplx ? 8 : 4;
        long long nb = c->svf_shape.bwd_blocks;
        const long long npair = K > 1 ? (long long)K * (K - 1) / 2 : 1;
        long long nchunk = nb * npair;
        long long cgs = c->svf_shape.fwd_cgs;
        if (c->q8 || c->q8c) {
            nb = std::max(nb, (long long)c->q8_shape.bwd_blocks);
            nchunk = std::max(nchunk, (long long)c->q8_shape.bwd_blocks);
            cgs = std::max(cgs, (long long)c->q8_shape.fwd_cgs);
            const long long rp = c->q8_shape.rpad;
            const size_t nsets = c->q8c ? 3 : 1;  // complex: digit sets X_A, X_R, X_I
            if (c->q8c) {
                HIPCHK(c, c->dw3.ensure((size_t)3 * S * 8), "alloc S*v sample sums");
                HIPCHK(c, c->dvin3.ensure((size_t)(3 * P + 4) * 8), "alloc S*v operands");
                HIPCHK(c, c->dsvout3.ensure((size_t)3 * (P + 1) * 8), "alloc S*v out");
            }
            HIPCHK(c, c->dUq.ensure(std::max((size_t)(nsets * rp * c->q8_shape.ucols * 3), c->q8_shape.ug_bytes)),
                   "alloc quantised factors");
            HIPCHK(c, c->descale.ensure((size_t)rp * 8), "alloc factor scales");
            HIPCHK(c, c->dUmax.ensure((size_t)rp * 4), "alloc factor row maxima");
            size_t c_vq = c->q8_shape.vq_bytes, c_q8w = c->q8_shape.w_bytes, c_part = 0, c_scr = 0;
            if (c->q8c) nq_q8c_sizes(c->q8_shape, P, &c_vq, &c_q8w, &c_part, &c_scr);  // three sets at once
            HIPCHK(c, c->dvq.ensure(c_vq), "alloc S*v operands");
            HIPCHK(c, c->dq8w.ensure(c_q8w), "alloc S*v weight digits");
            HIPCHK(c, c->dsvfpart.ensure(c_part), "alloc S*v partials");
            HIPCHK(c, c->dsvfscratch.ensure(c_scr), "alloc S*v scratch");
            HIPCHK(c, c->dq8m.ensure(c->q8_shape.m_bytes), "alloc spin masks");
            if (c->q8_shape.gemm)
                HIPCHK(c, c->dUf.ensure(nsets * c->q8_shape.uf_bytes), "alloc row-major factor digits");
            if (c->q8_shape.xc_bytes)
                HIPCHK(c, c->dxc.ensure(nsets * c->q8_shape.xc_bytes), "alloc factor c columns");
        }
        HIPCHK(c, c->dU.ensure((size_t)S * K * c->svf_shape.ucols * tsz), "alloc S*v factors");
        HIPCHK(c, c->dwrow.ensure((size_t)cgs * S * K * 16), "alloc S*v rows");
        HIPCHK(c, c->dsvfpart.ensure((size_t)std::max(nb * 16, nchunk * 8) * (P + 1)), "alloc S*v partials");
        HIPCHK(c, c->dsvfscratch.ensure((size_t)(nchunk / 32 + 2) * (P + 1) * 16), "alloc S*v scratch");
        HIPCHK(c, c->dsvftmp.ensure((size_t)3 * (P + 1) * 16), "alloc moments");
    }
    const size_t obytes =
        c->svf ? 16 : (size_t)S * c->Pld * o_elem(c->prec, c->prec == NQES_PREC_FP64 ? !c->oreal : cplx);
    size_t freeb = 0, totalb = 0;
    HIPCHK(c, hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
    long long need = (long long)obytes - (long long)c->dO.bytes;
    if (need > 0 && (size_t)need + (size_t)(1ull << 30) > freeb)
        return c->fail(NQES_ENOMEM, "O needs %.2f GB, %.2f GB free", obytes / 1e9, freeb / 1e9);
    HIPCHK(c, c->dO.ensure(obytes), "alloc O");
    HIPCHK(c, c->dHs.ensure((size_t)(S * K * K * 16)), "alloc H");
    HIPCHK(c, c->dEs.ensure((size_t)(S * 16)), "alloc E");
    HIPCHK(c, c->dspins.ensure((size_t)(S * K * c->NW * 4)), "alloc spins");
    HIPCHK(c, c->dacc.ensure((size_t)C * 8), "alloc accepts");
    HIPCHK(c, c->dties.ensure((size_t)C * 8), "alloc tie counts");
    if (c->prec != NQES_PREC_FP64) HIPCHK(c, hipMemsetAsync(c->dties.p, 0, (size_t)C * 8, c->stream), "clear ties");
    c->mchunks = choose_chunks(S, P, 256);
    const size_t mpart = c->svf ? 16 : (size_t)c->mchunks * P * 16;  // factored: moments from the factors
    HIPCHK(c, c->dmpart_o.ensure(mpart), "alloc partials");
    HIPCHK(c, c->dmpart_s.ensure(mpart), "alloc partials");
    HIPCHK(c, c->dmpart_f.ensure(mpart), "alloc partials");
    HIPCHK(c, c->dmo.ensure((size_t)P * 16), "alloc moments");
    HIPCHK(c, c->dms.ensure((size_t)P * 16), "alloc moments");
    HIPCHK(c, c->dmf.ensure((size_t)P * 16), "alloc moments");
    int hch = choose_chunks(S, K * K, 64);
    HIPCHK(c, c->dhpart.ensure((size_t)hch * K * K * 16), "alloc partials");
    HIPCHK(c, c->dhsum.ensure((size_t)K * K * 16), "alloc partials");
    int ech = choose_chunks(S, 1, 64);
    HIPCHK(c, c->depart.ensure((size_t)ech * 16), "alloc partials");
    HIPCHK(c, c->desum.ensure(16), "alloc partials");

    HIPCHK(c, hipMemcpyAsync(c->dW.p, W, 16 * P, hipMemcpyHostToDevice, c->stream), "copy W");
    int st = (c->prec == NQES_PREC_FP64) ? sample_fp64(c, seed, C, m, B) : sample_mixed(c, seed, C, m, B);
    if (st != NQES_OK) return st;
    HIPCHK(c, nq_launch_colsum64(c->dHs.p, S, K * K, hch, c->dhpart.p, c->dhsum.p, c->stream), "colsum kernel");
    HIPCHK(c, nq_launch_colsum64(c->dEs.p, S, 1, ech, c->depart.p, c->desum.p, c->stream), "colsum kernel");
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");

    std::vector<cd> so(P), ss(P), sf(P), sh(K * K), se(1);
    unsigned int flags = 0;
    HIPCHK(c, hipMemcpyAsync(so.data(), c->dmo.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(ss.data(), c->dms.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(sf.data(), c->dmf.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(sh.data(), c->dhsum.p, 16 * K * K, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(se.data(), c->desum.p, 16, hipMemcpyDeviceToHost, c->stream), "copy moments");
    if (c->prec == NQES_PREC_MIXED)
        HIPCHK(c, hipMemcpyAsync(&flags, c->dflags.p, 4, hipMemcpyDeviceToHost, c->stream), "copy flags");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sample sync");
    if (c->prec == NQES_PREC_MIXED) tune_record(c, event_ms(c->ev[0], c->ev[1]));

    // normalisation (rbm.cc:641-650)
    using namespace nq64;
    const double Sd = (double)S;
    c->eloc = divr(se[0], Sd);
    c->o.resize(P);
    c->skk.resize(P);
    c->f.resize(P);
    int nonfinite = !(std::isfinite(c->eloc.re) && std::isfinite(c->eloc.im)) || (flags & 1u);
    for (long long q = 0; q < P; q++) {
        cd oq = divr(so[q], Sd);
        cd sq = divr(ss[q], Sd);
        sq.re -= norm(oq);
        sq = scale(reglam, sq);
        cd fq = divr(sf[q], Sd);
        fq = sub(fq, mul(c->eloc, conj(oq)));
        c->o[q] = oq;
        c->skk[q] = sq;
        c->f[q] = fq;
        if (!(std::isfinite(fq.re) && std::isfinite(fq.im) && std::isfinite(sq.re))) nonfinite = 1;
    }
    c->emat.resize(K * K);
    for (int q = 0; q < K * K; q++) c->emat[q] = divr(sh[q], Sd);
    c->nonfinite = nonfinite;
    c->theta_overflow = (flags & 1u) ? 1 : 0;
    c->C = C;
    c->m = m;
    c->B = B;
    c->S = S;
    HIPCHK(c, c->dv.ensure((size_t)(P + 1) * 16), "alloc v");
    HIPCHK(c, c->dov.ensure(16), "alloc ov");
    HIPCHK(c, c->dw.ensure((size_t)S * 16), "alloc w");
    if (c->prec == NQES_PREC_FP64 && c->sv64.ok) {  // one pass: per-block partial rows
        const long long nvec = c->oreal ? c->Pld / 2 : P;
        const long long nb = c->sv64.blocks;
        HIPCHK(c, c->dsvpart.ensure((size_t)nb * nvec * 16), "alloc S*v partials");
        HIPCHK(c, c->dsvwsum.ensure((size_t)nb * 16), "alloc S*v partials");
        HIPCHK(c, c->dsvscratch.ensure((size_t)(nb / 32 + 2) * (P + 1) * 16), "alloc S*v scratch");
    } else if (c->prec == NQES_PREC_FP64) {
        c->svchunks = choose_chunks(S, P, 256);
        HIPCHK(c, c->dsvpart.ensure((size_t)c->svchunks * P * 16), "alloc S*v partials");
    } else if (!c->svf) {
        const long long nvec = c->Pld / (cplx ? 2 : 4);
        const SvShape sh = nq_sv_shape(cplx, nvec);
        long long nblk = sh.qv ? sh.blocks : choose_chunks(S, P, 256);  // one pass / two passes
        if (!sh.qv && nblk > S) nblk = S;
        c->svchunks = (int)nblk;
        HIPCHK(c, c->dsvpart.ensure((size_t)c->svchunks * c->Pld * 16), "alloc S*v partials");
        HIPCHK(c, c->dsvwpart.ensure((size_t)c->svchunks * 16), "alloc S*v partials");
        HIPCHK(c, c->dsvscratch.ensure((size_t)(c->svchunks / 32 + 2) * (c->Pld + 1) * 16), "alloc S*v scratch");
    }
    HIPCHK(c, c->dsvout.ensure((size_t)(P + 1) * 16), "alloc S*v out");
    c->svraw.resize(P + 1);
    c->sv_calls = 0;
    c->sv_ms = 0;
    if (c->dense) {  // the dense S of rbm.cc:651-657 from the stored O rows
        HIPCHK(c, c->doo.ensure((size_t)P * P * 16), "alloc dense S");
        HIPCHK(c, c->dvo.ensure((size_t)P * 16), "alloc o");
        HIPCHK(c, c->dskk.ensure((size_t)P * 16), "alloc skk");
        HIPCHK(c, hipMemcpyAsync(c->dvo.p, c->o.data(), 16 * P, hipMemcpyHostToDevice, c->stream), "copy o");
        HIPCHK(c, hipMemcpyAsync(c->dskk.p, c->skk.data(), 16 * P, hipMemcpyHostToDevice, c->stream), "copy skk");
        HIPCHK(c, nq_launch_dense_gram64(c->dO.p, S, P, c->dvo.p, c->doo.p, c->stream), "dense S kernel");
        HIPCHK(c, hipStreamSynchronize(c->stream), "dense S sync");
    }
    c->sampled = true;
    return NQES_OK;
}

int nqes_enumerate(nqes_ctx *c, const double *W, int64_t cfg0, int64_t ncfg, double *phi, double *res,
                   double *tanh_)
{
    if (!c) return NQES_EINVAL;
    if (!c->has_problem) return c->fail(NQES_ESTATE, "no problem: call nqes_problem_init first");
    if (!W) return c->fail(NQES_EINVAL, "W is NULL");
    const int N = c->N, M = c->M, K = c->K;
    if (N * K >= 32) return c->fail(NQES_EINVAL, "N*K = %d: exhaustive evaluation needs N*K < 32", N * K);
    if (cfg0 < 0 || ncfg < 0 || cfg0 + ncfg > (1ll << (N * K)) || ncfg > (1ll << 24))
        return c->fail(NQES_EINVAL, "configurations [%lld, %lld) outside [0, 2^%d) or more than 2^24 at once",
                       (long long)cfg0, (long long)(cfg0 + ncfg), N * K);
    if (ncfg == 0) return NQES_OK;
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    const size_t kk = (size_t)K * K, nphi = 16 * kk * ncfg, ntanh = 16 * kk * M * ncfg;
    HIPCHK(c, c->denum.ensure(2 * nphi + ntanh), "alloc evaluation outputs");
    char *base = static_cast<char *>(c->denum.p);
    HIPCHK(c, hipMemcpyAsync(c->dW.p, W, 16 * c->P, hipMemcpyHostToDevice, c->stream), "copy W");
    HIPCHK(c, nq_launch_tables64(N, M, K, c->L, c->dW.p, c->dchT.p, c->dshT.p, c->deaP.p, c->deaN.p, c->stream),
           "tables kernel");
    Fp64SampleArgs a{};
    a.N = N;
    a.M = M;
    a.NW = c->NW;
    a.L = c->L;
    a.P = c->P;
    a.h = c->h;
    a.model = c->model;
    a.delta = c->model == NQES_MODEL_XXZ ? -1.0 : 1.0;
    a.J = static_cast<const double *>(c->dJ.p);
    a.W = c->dW.p;
    a.chT = c->dchT.p;
    a.shT = c->dshT.p;
    a.eaP = c->deaP.p;
    a.eaN = c->deaN.p;
    HIPCHK(c, nq_launch_enum64(K, a, cfg0, ncfg, base, base + nphi, base + 2 * nphi, c->stream), "evaluation kernel");
    if (phi) HIPCHK(c, hipMemcpyAsync(phi, base, nphi, hipMemcpyDeviceToHost, c->stream), "copy Phi");
    if (res) HIPCHK(c, hipMemcpyAsync(res, base + nphi, nphi, hipMemcpyDeviceToHost, c->stream), "copy res");
    if (tanh_) HIPCHK(c, hipMemcpyAsync(tanh_, base + 2 * nphi, ntanh, hipMemcpyDeviceToHost, c->stream), "copy tanh");
    HIPCHK(c, hipStreamSynchronize(c->stream), "evaluation sync");
    c->sampled = false;  // the tables now hold W's
    return NQES_OK;
}

int nqes_set_dense_S(nqes_ctx *c, int dense)
{
    if (!c) return NQES_EINVAL;
    if (!c->has_problem) return c->fail(NQES_ESTATE, "no problem: call nqes_problem_init first");
    if (dense && c->prec != NQES_PREC_FP64) return c->fail(NQES_EINVAL, "the dense S form needs NQES_PREC_FP64");
    c->dense = dense != 0;
    c->sampled = false;
    return NQES_OK;
}

int nqes_get_samples(nqes_ctx *c, uint32_t *spins, double *H, uint64_t *accepts)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    if (spins)
        HIPCHK(c, copy_sync(spins, c->dspins.p, (size_t)c->S * c->K * c->NW * 4, hipMemcpyDeviceToHost), "copy spins");
    if (H) HIPCHK(c, copy_sync(H, c->dHs.p, (size_t)c->S * c->K * c->K * 16, hipMemcpyDeviceToHost), "copy H");
    if (accepts) HIPCHK(c, copy_sync(accepts, c->dacc.p, (size_t)c->C * 8, hipMemcpyDeviceToHost), "copy accepts");
    return NQES_OK;
}

// O rows rebuilt from the factor rows and spins with the FP32 additions (in replica
// order) of the measurement kernel; they equal the rows the streamed path stores up to
// the FP32 rounding of c * tanh(theta), which the kernel may contract into an FMA:
// a_i = sum_kk +-c_{k,kk}, b_j = sum_kk U_kk[kM+j], w_ij = sum_kk +-U_kk[kM+j]
static int get_O_factored(nqes_ctx *c, int64_t row0, int64_t nrows, double *out)
{
    const int N = c->N, M = c->M, K = c->K, NW = c->NW, UC = c->svf_shape.ucols, L = c->L;
    const size_t tsz = c->cplx ? 8 : 4;
    std::vector<float> U((size_t)nrows * K * UC * (c->cplx ? 2 : 1));
    std::vector<uint32_t> sp((size_t)nrows * K * NW);
    HIPCHK(c, copy_sync(U.data(), static_cast<char *>(c->dU.p) + (size_t)row0 * K * UC * tsz, U.size() * 4,
                        hipMemcpyDeviceToHost),
           "copy S*v factors");
    HIPCHK(c, copy_sync(sp.data(), static_cast<uint32_t *>(c->dspins.p) + (size_t)row0 * K * NW, sp.size() * 4,
                        hipMemcpyDeviceToHost),
           "copy spins");
    const int R = c->cplx ? 2 : 1;
    for (int64_t r = 0; r < nrows; r++) {
        auto u = [&](int kk, int col, int part) { return U[(((size_t)r * K + kk) * UC + col) * R + part]; };
        auto bit = [&](int kk, int i) { return (sp[((size_t)r * K + kk) * NW + (i >> 5)] >> (i & 31)) & 1u; };
        double *row = out + 2 * (size_t)r * c->P;
        for (int k = 0; k < K; k++)
            for (int part = 0; part < R; part++) {
                for (int i = 0; i < N; i++) {
                    float v = 0.0f;
                    for (int kk = 0; kk < K; kk++) v = bit(kk, i) ? v - u(kk, K * M + k, part) : v + u(kk, K * M + k, part);
                    row[2 * ((size_t)k * L + i) + part] = v;
                }
                for (int j = 0; j < M; j++) {
                    float v = 0.0f;
                    for (int kk = 0; kk < K; kk++) v = v + u(kk, k * M + j, part);
                    row[2 * ((size_t)k * L + N + j) + part] = v;
                    for (int i = 0; i < N; i++) {
                        float w = 0.0f;
                        for (int kk = 0; kk < K; kk++) w = bit(kk, i) ? w - u(kk, k * M + j, part) : w + u(kk, k * M + j, part);
                        row[2 * ((size_t)k * L + N + M + (size_t)i * M + j) + part] = w;
                    }
                }
            }
        if (R == 1)
            for (long long q = 0; q < c->P; q++) row[2 * q + 1] = 0.0;
    }
    return NQES_OK;
}

int nqes_get_O(nqes_ctx *c, int64_t row0, int64_t nrows, double *out)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    if (!out || row0 < 0 || nrows < 0 || row0 + nrows > c->S) return c->fail(NQES_EINVAL, "row range out of bounds");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    if (c->svf) return get_O_factored(c, row0, nrows, out);
    const bool fp64 = c->prec == NQES_PREC_FP64;
    const size_t e = o_elem(c->prec, fp64 ? !c->oreal : c->cplx);
    std::vector<unsigned char> buf((size_t)nrows * c->Pld * e);
    HIPCHK(c,
           copy_sync(buf.data(), static_cast<char *>(c->dO.p) + (size_t)row0 * c->Pld * e, buf.size(),
                     hipMemcpyDeviceToHost),
           "copy O");
    for (int64_t r = 0; r < nrows; r++)
        for (long long q = 0; q < c->P; q++) {
            const unsigned char *src = buf.data() + ((size_t)r * c->Pld + q) * e;
            double *dst = out + 2 * ((size_t)r * c->P + q);
            if (e == 16) {
                memcpy(dst, src, 16);
            } else if (fp64) {  // real weights: Re O stored
                memcpy(dst, src, 8);
                dst[1] = 0.0;
            } else if (e == 8) {
                float v[2];
                memcpy(v, src, 8);
                dst[0] = v[0];
                dst[1] = v[1];
            } else {
                float v;
                memcpy(&v, src, 4);
                dst[0] = v;
                dst[1] = 0.0;
            }
        }
    return NQES_OK;
}

int nqes_get_ulp_ties(nqes_ctx *c, uint64_t *per_chain, uint64_t *total)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    std::vector<uint64_t> t((size_t)c->C);
    HIPCHK(c, copy_sync(t.data(), c->dties.p, (size_t)c->C * 8, hipMemcpyDeviceToHost), "copy tie counts");
    if (per_chain) memcpy(per_chain, t.data(), (size_t)c->C * 8);
    if (total) {
        uint64_t s = 0;
        for (auto v : t) s += v;
        *total = s;
    }
    return NQES_OK;
}

int nqes_get_sampler_launch(nqes_ctx *c, int *chains_per_block, int *tables_in_lds, int *samples_per_block,
                            int *tuning)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled || c->prec != NQES_PREC_MIXED)
        return c->fail(NQES_ESTATE, "no mixed-precision samples: call nqes_sample first");
    const SweepShape sw = nq_sweep_shape(c->cplx, c->N, c->M, c->K, c->lds_max);
    const MeasureShape ms = nq_measure_shape(c->cplx, c->N, c->K, sw, c->lds_max);
    if (chains_per_block) *chains_per_block = c->launch.sweep_wpb * (64 / sw.G);
    if (tables_in_lds) *tables_in_lds = c->launch.sweep_tab;
    if (samples_per_block) *samples_per_block = c->launch.measure_wpb * (64 / ms.G);
    if (tuning) *tuning = c->tune_state;
    return NQES_OK;
}

int nqes_get_moments(nqes_ctx *c, double *o, double *skk, double *f, double *eloc, double *emat)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    if (o) memcpy(o, c->o.data(), 16 * c->P);
    if (skk) memcpy(skk, c->skk.data(), 16 * c->P);
    if (f) memcpy(f, c->f.data(), 16 * c->P);
    if (eloc) memcpy(eloc, &c->eloc, 16);
    if (emat) memcpy(emat, c->emat.data(), 16 * c->K * c->K);
    return NQES_OK;
}

// the int8 S*v of the current sample set as one graph: v and o^T v from pinned host
// memory, the kernels of nq_q8_apply, the raw result back to pinned memory (no memset
// nodes).  NQES_GRAPH=0 launches the same sequence eagerly instead.
// doubles in and out of one graphed S*v: real weights v | o^T v in and the P + 1 raw
// sums out; complex weights (q8c) v_A | v_R | v_I | o^T v | 0, 0 in and the three
// backward runs' P + 1 raw sums out
static long long sv_in_len(const nqes_ctx *c) { return c->q8c ? 3 * c->P + 4 : c->P + 1; }
static long long sv_out_len(const nqes_ctx *c) { return c->q8c ? 3 * (c->P + 1) : c->P + 1; }

// complex weights: the raw sums sum_s conj(O_s) w_s (p < P) and sum_s w_s (p = P) from
// the three backward runs A, R, I (Gauss's form): Re = A - R - I, Im = R - I; at P the
// runs' weight sums, Re w over set I and Im w over set R
static void q8c_raw(nqes_ctx *c, const double *r)
{
    const long long P = c->P;
    const double *ra = r, *rr = r + (P + 1), *ri = r + 2 * (P + 1);
    for (long long q = 0; q < P; q++) c->svraw[q] = nq64::mk((ra[q] - rr[q]) - ri[q], rr[q] - ri[q]);
    c->svraw[P] = nq64::mk(ri[P], rr[P]);
}

// the int8 S*v sequence on the context's stream: v (and o^T v) from pinned memory, the
// kernels, the raw result back to pinned memory
static hipError_t enqueue_sv(nqes_ctx *c)
{
    const long long P = c->P;
    const long long nin = sv_in_len(c), nout = sv_out_len(c);
    // v and o^T v arrive in one copy: the kernels read o^T v right after v (the v buffer
    // holds 16 (P + 1) bytes, room for 2 P + 2 doubles)
    SvfArgs f = svf_args(c, c->S);
    f.ov = (const double *)c->dv.p + P;
    void *din = c->q8c ? c->dvin3.p : c->dv.p, *dout = c->q8c ? c->dsvout3.p : c->dsvout.p;
    hipError_t e = hipMemcpyAsync(din, c->hv_pin, 8 * nin, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = c->q8c ? nq_q8c_apply(q8c_args(c), c->q8_shape, static_cast<const double *>(din),
                                  static_cast<double *>(c->dw3.p), c->dsvfscratch.p, static_cast<double *>(dout),
                                  c->stream)
                   : nq_q8_apply(f, c->q8_shape, c->dsvfscratch.p, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->hr_pin, dout, 8 * nout, hipMemcpyDeviceToHost, c->stream);
    return e;
}

// pinned staging for the S*v input and result
static int ensure_sv_pinned(nqes_ctx *c)
{
    const long long npin = std::max(sv_in_len(c), sv_out_len(c));
    if ((size_t)npin <= c->pin_len) return NQES_OK;
    NqUnsafeHold hold;
    if (c->hv_pin) (void)hipHostFree(c->hv_pin);
    if (c->hr_pin) (void)hipHostFree(c->hr_pin);
    c->hv_pin = c->hr_pin = nullptr;
    c->pin_len = 0;
    HIPCHK(c, hipHostMalloc((void **)&c->hv_pin, 8 * npin, hipHostMallocDefault), "pinned v");
    HIPCHK(c, hipHostMalloc((void **)&c->hr_pin, 8 * npin, hipHostMallocDefault), "pinned result");
    c->pin_len = (size_t)npin;
    return NQES_OK;
}

// Captures enqueue_sv into a graph (c->svx).  The capture holds nq_capture_mutex()
// exclusively, so this library's allocations and synchronous copies on other threads
// wait for it; a capture that fails anyway (the runtime can end a capture because of a
// call made on another thread) leaves c->svx empty, and the caller launches the same
// sequence directly, as the graph would, and captures again on its next call.
static void build_sv_graph(nqes_ctx *c)
{
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    {
        std::unique_lock<std::shared_mutex> lk(nq_capture_mutex());
        nq_capturing() = true;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            e = enqueue_sv(c);
            const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (e == hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        nq_capturing() = false;
        if (e != hipSuccess) {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return;
        }
    }
    c->svg = g;
    c->svx = x;
}

// the int8 S*v graph in two halves: launch (the sv_in_len doubles of `in` to pinned
// memory, graph on the stream) and finish (wait, raw result to svraw); the host may
// work in between
static int sv_graph_launch(nqes_ctx *c, const double *in)
{
    int st = ensure_sv_pinned(c);
    if (st != NQES_OK) return st;
    if (!c->svx) build_sv_graph(c);
    memcpy(c->hv_pin, in, 8 * sv_in_len(c));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
    if (c->svx) HIPCHK(c, hipGraphLaunch(c->svx, c->stream), "S*v graph launch");
    else HIPCHK(c, enqueue_sv(c), "S*v kernels");
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
    return NQES_OK;
}

static int sv_graph_finish(nqes_ctx *c)
{
    HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
    const long long P = c->P;
    if (c->q8c) q8c_raw(c, c->hr_pin);
    else memcpy(c->svraw.data(), c->hr_pin, 8 * (P + 1));
    c->sv_ms += event_ms(c->ev[2], c->ev[3]);
    c->sv_calls++;
    return NQES_OK;
}

static bool sv_graph_on(const nqes_ctx *c) { return (c->q8 || c->q8c) && env_int("NQES_GRAPH", 1) != 0; }

// device part of S*v on the FP32 O of the mixed path: raw[p] = sum_s conj(O_sp) w_s,
// raw[P] = sum_s w_s, w_s = O_s^T v - o^T v.  v given as D = double (real) or cd.
static int sv_mixed_device(nqes_ctx *c, const void *v, const void *ov, size_t dsize)
{
    const long long P = c->P;
    if (sv_graph_on(c)) {  // real weights, int8 form: v (P doubles) then o^T v
        c->hv.resize(P + 1);
        if (v != c->hv.data()) memcpy(c->hv.data(), v, 8 * P);
        c->hv[P] = *(const double *)ov;
        int st = sv_graph_launch(c, c->hv.data());
        return st != NQES_OK ? st : sv_graph_finish(c);
    }
    HIPCHK(c, hipMemcpyAsync(c->dv.p, v, dsize * P, hipMemcpyHostToDevice, c->stream), "copy v");
    HIPCHK(c, hipMemcpyAsync(c->dov.p, ov, dsize, hipMemcpyHostToDevice, c->stream), "copy ov");
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
    if (c->q8) {
        const SvfArgs f = svf_args(c, c->S);
        HIPCHK(c, nq_q8_apply(f, c->q8_shape, c->dsvfscratch.p, c->dsvout.p, c->stream), "int8 factored S*v kernels");
    } else if (c->svf) {
        const SvfArgs f = svf_args(c, c->S);
        HIPCHK(c, nq_launch_svf(c->cplx, f, c->svf_shape, c->dsvfscratch.p, c->dsvout.p, c->stream),
               "factored S*v kernels");
    } else {
        HIPCHK(c, nq_launch_sv_mx(c->cplx, c->dO.p, c->dv.p, c->dov.p, c->S, P, c->Pld, c->svchunks, c->dsvpart.p,
                                  c->dsvwpart.p, c->dw.p, c->dsvscratch.p, c->dsvout.p, c->stream),
               "S*v kernels");
    }
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
    HIPCHK(c, hipMemcpyAsync(c->svraw.data(), c->dsvout.p, dsize * (P + 1), hipMemcpyDeviceToHost, c->stream),
           "copy S*v");
    HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
    c->sv_ms += event_ms(c->ev[2], c->ev[3]);
    c->sv_calls++;
    return NQES_OK;
}

// device part of the one-pass FP64 S*v: raw sums over the stored O rows (svraw), v
// given as D = double (real O) or cd; v and o^T v go to the device in one copy
static int sv_fp64_device(nqes_ctx *c, const void *v, const void *ov, size_t dsize)
{
    const long long P = c->P;
    char *dv = static_cast<char *>(c->dv.p);
    HIPCHK(c, hipMemcpyAsync(dv, v, dsize * P, hipMemcpyHostToDevice, c->stream), "copy v");
    HIPCHK(c, hipMemcpyAsync(dv + dsize * P, ov, dsize, hipMemcpyHostToDevice, c->stream), "copy ov");
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
    const long long nvec = c->oreal ? c->Pld / 2 : P;
    HIPCHK(c, nq_launch_sv_rows64(!c->oreal, c->sv64, c->dO.p, dv, dv + dsize * P, c->S, nvec, P, c->dsvpart.p,
                                  c->dsvwsum.p, c->dsvscratch.p, c->dsvout.p, c->stream),
           "S*v kernels");
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
    HIPCHK(c, hipMemcpyAsync(c->svraw.data(), c->dsvout.p, dsize * (P + 1), hipMemcpyDeviceToHost, c->stream),
           "copy S*v");
    HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
    c->sv_ms += event_ms(c->ev[2], c->ev[3]);
    c->sv_calls++;
    return NQES_OK;
}

// the complex-weight S*v's operands into hv (the graphed and the eager form read it
// alike): v_A = Re v, v_R = Im v - Re v, v_I = Re v + Im v, o^T v, two zeros
static void q8c_prepare(nqes_ctx *c, const cd *v, cd ov)
{
    const long long P = c->P;
    c->hv.resize(3 * P + 4);
    for (long long q = 0; q < P; q++) {
        c->hv[q] = v[q].re;
        c->hv[P + q] = v[q].im - v[q].re;
        c->hv[2 * P + q] = v[q].re + v[q].im;
    }
    c->hv[3 * P] = ov.re;
    c->hv[3 * P + 1] = ov.im;
    c->hv[3 * P + 2] = 0.0;
    c->hv[3 * P + 3] = 0.0;
}

static int sv_q8c_device(nqes_ctx *c, const cd *v, cd ov)
{
    const long long P = c->P;
    q8c_prepare(c, v, ov);
    if (sv_graph_on(c)) {
        int st = sv_graph_launch(c, c->hv.data());
        return st != NQES_OK ? st : sv_graph_finish(c);
    }
    double *din = static_cast<double *>(c->dvin3.p);
    HIPCHK(c, hipMemcpyAsync(din, c->hv.data(), 8 * (3 * P + 4), hipMemcpyHostToDevice, c->stream), "copy v");
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
    HIPCHK(c, nq_q8c_apply(q8c_args(c), c->q8_shape, din, static_cast<double *>(c->dw3.p), c->dsvfscratch.p,
                           static_cast<double *>(c->dsvout3.p), c->stream),
           "complex int8 factored S*v kernels");
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
    std::vector<double> raw(3 * (P + 1));
    HIPCHK(c, hipMemcpyAsync(raw.data(), c->dsvout3.p, 8 * 3 * (P + 1), hipMemcpyDeviceToHost, c->stream), "copy S*v");
    HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
    q8c_raw(c, raw.data());
    c->sv_ms += event_ms(c->ev[2], c->ev[3]);
    c->sv_calls++;
    return NQES_OK;
}

// real O and o (real weights, either path): part 0 / 1 of v (real / imaginary) into hv,
// returns o^T hv
static double mixed_real_prepare(nqes_ctx *c, const cd *v, int part)
{
    const long long P = c->P;
    c->hv.resize(P + 1);
    double ovp = 0.0;
    for (long long q = 0; q < P; q++) {
        c->hv[q] = part ? v[q].im : v[q].re;
        ovp += c->o[q].re * c->hv[q];
    }
    return ovp;
}

// (1/S) [raw - o sum_s w_s] + skk .* hv into part 0 (imaginary part cleared) or part 1 of out
static void mixed_real_result(nqes_ctx *c, cd *out, int part)
{
    const long long P = c->P;
    const double Sd = (double)c->S;
    const double *raw = reinterpret_cast<const double *>(c->svraw.data());
    const double wsum = raw[P];
    for (long long q = 0; q < P; q++) {
        double r = (raw[q] - c->o[q].re * wsum) / Sd + c->skk[q].re * c->hv[q];
        if (part == 0) {
            out[q].re = r;
            out[q].im = 0.0;
        } else {
            out[q].im = r;
        }
    }
}

// complex O: out = (1/S) [raw - conj(o) raw[P]] + skk .* v from the raw sums in svraw
static void cplx_result(nqes_ctx *c, const cd *v, cd *out)
{
    using namespace nq64;
    const long long P = c->P;
    const double Sd = (double)c->S;
    const cd wsum = c->svraw[P];
    for (long long q = 0; q < P; q++) {
        cd r = divr(sub(c->svraw[q], mul(conj(c->o[q]), wsum)), Sd);
        out[q] = add(r, mul(c->skk[q], v[q]));
    }
}

// S v = (1/S) sum_s (O_s^T v) conj(O_s) + skk .* v - (o^T v) conj(o)   (rbm.cc:661-684)
static int apply_S(nqes_ctx *c, const cd *v, cd *out)
{
    using namespace nq64;
    const long long P = c->P;
    const double Sd = (double)c->S;
    cd ov = zero();
    for (long long q = 0; q < P; q++) ov = add(ov, mul(c->o[q], v[q]));
    if (c->dense) {  // rbm.cc:686-696
        HIPCHK(c, hipMemcpyAsync(c->dv.p, v, 16 * P, hipMemcpyHostToDevice, c->stream), "copy v");
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
        HIPCHK(c, nq_launch_dense_apply64(c->doo.p, c->dskk.p, c->dv.p, P, c->dsvout.p, c->stream), "dense S kernel");
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
        HIPCHK(c, hipMemcpyAsync(out, c->dsvout.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy S*v");
        HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
        c->sv_ms += event_ms(c->ev[2], c->ev[3]);
        c->sv_calls++;
        return NQES_OK;
    }
    if (c->prec == NQES_PREC_FP64 && c->sv64.ok) {
        if (!c->oreal) {  // complex O: (1/S) [sum_s conj(O_s) w_s - conj(o) sum_s w_s] + skk .* v
            int st = sv_fp64_device(c, v, &ov, 16);
            if (st != NQES_OK) return st;
            const cd wsum = c->svraw[P];
            for (long long q = 0; q < P; q++) {
                cd r = divr(sub(c->svraw[q], mul(conj(c->o[q]), wsum)), Sd);
                out[q] = add(r, mul(c->skk[q], v[q]));
            }
            return NQES_OK;
        }
        bool has_im = false;  // real O: the real and imaginary parts of v separately
        for (long long q = 0; q < P; q++)
            if (v[q].im != 0.0) has_im = true;
        for (int part = 0; part < (has_im ? 2 : 1); part++) {
            double ovp = mixed_real_prepare(c, v, part);
            int st = sv_fp64_device(c, c->hv.data(), &ovp, 8);
            if (st != NQES_OK) return st;
            mixed_real_result(c, out, part);
        }
        return NQES_OK;
    }
    if (c->prec == NQES_PREC_FP64) {
        HIPCHK(c, hipMemcpyAsync(c->dv.p, v, 16 * P, hipMemcpyHostToDevice, c->stream), "copy v");
        HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
        HIPCHK(c,
               nq_launch_sv64(c->dO.p, c->dv.p, c->S, P, c->svchunks, c->dw.p, c->dsvpart.p, c->dsvout.p, c->stream),
               "S*v kernels");
        HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
        HIPCHK(c, hipMemcpyAsync(c->svraw.data(), c->dsvout.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy S*v");
        HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
        c->sv_ms += event_ms(c->ev[2], c->ev[3]);
        c->sv_calls++;
        for (long long q = 0; q < P; q++) {
            cd r = divr(c->svraw[q], Sd);
            out[q] = add(r, sub(mul(c->skk[q], v[q]), mul(ov, conj(c->o[q]))));
        }
        return NQES_OK;
    }
    // mixed: centred single pass, (1/S) [sum_s conj(O_s) w_s - conj(o) sum_s w_s]
    if (c->cplx) {
        int st = c->q8c ? sv_q8c_device(c, v, ov) : sv_mixed_device(c, v, &ov, 16);
        if (st != NQES_OK) return st;
        cplx_result(c, v, out);
        return NQES_OK;
    }
    // real O and o: apply to the real and imaginary parts of v separately
    bool has_im = false;
    for (long long q = 0; q < P; q++)
        if (v[q].im != 0.0) has_im = true;
    for (int part = 0; part < (has_im ? 2 : 1); part++) {
        double ovp = mixed_real_prepare(c, v, part);
        int st = sv_mixed_device(c, c->hv.data(), &ovp, 8);
        if (st != NQES_OK) return st;
        mixed_real_result(c, out, part);
    }
    return NQES_OK;
}

// S*v as launch and finish for the solver: the int8 graph path with a real v (the
// real-weight solve) returns from the launch while the device works; every other form
// runs whole in the finish.  Results equal apply_S's bit for bit.
struct SvOperator {
    nqes_ctx *c;
    int status = NQES_OK;
    bool split = false;
    const std::vector<cd> *v = nullptr;
    bool launch(const std::vector<cd> &vv)
    {
        v = &vv;
        split = false;
        if (c->dense || c->prec != NQES_PREC_MIXED || !sv_graph_on(c)) return true;
        const long long P = c->P;
        if (c->q8c) {  // complex weights: Re v | Im v | o^T v
            cd ov = nq64::zero();
            for (long long q = 0; q < P; q++) ov = nq64::add(ov, nq64::mul(c->o[q], vv[q]));
            q8c_prepare(c, vv.data(), ov);
            status = sv_graph_launch(c, c->hv.data());
            split = status == NQES_OK;
            return split;
        }
        for (long long q = 0; q < P; q++)
            if (vv[q].im != 0.0) return true;
        c->hv.resize(P + 1);
        c->hv[P] = mixed_real_prepare(c, vv.data(), 0);
        status = sv_graph_launch(c, c->hv.data());
        split = status == NQES_OK;
        return split;
    }
    bool finish(std::vector<cd> &out)
    {
        if (!split) {
            status = apply_S(c, v->data(), out.data());
            return status == NQES_OK;
        }
        split = false;
        status = sv_graph_finish(c);
        if (status != NQES_OK) return false;
        if (c->q8c) cplx_result(c, v->data(), out.data());
        else mixed_real_result(c, out.data(), 0);
        return true;
    }
};

int nqes_apply_S(nqes_ctx *c, const double *v, double *out)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    if (!v || !out) return c->fail(NQES_EINVAL, "NULL vector");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    return apply_S(c, reinterpret_cast<const cd *>(v), reinterpret_cast<cd *>(out));
}

static int solve_impl(nqes_ctx *c, double rtol, uint64_t maxit, std::vector<cd> &x, nqmq::Result &res)
{
    bool aborted = false;
    SvOperator op{c};
    res = nqmq::solve_async(op, c->f, x, rtol, maxit, aborted);
    return aborted ? op.status : NQES_OK;
}

int nqes_solve(nqes_ctx *c, double rtol, uint64_t maxit, double *x, int *flag, uint64_t *iters)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    std::vector<cd> xv;
    nqmq::Result r;
    int st = solve_impl(c, rtol, maxit, xv, r);
    if (st != NQES_OK) return st;
    if (x) memcpy(x, xv.data(), 16 * c->P);
    if (flag) *flag = r.flag;
    if (iters) *iters = r.iters;
    return NQES_OK;
}

int nqes_correlations(nqes_ctx *c, const double *W, uint64_t seed, int C, int m, int B, double *x, double *z,
                      double *xx, double *zz)
{
    if (!c) return NQES_EINVAL;
    if (!c->has_problem) return c->fail(NQES_ESTATE, "nqes_problem_init not called");
    if (!W) return c->fail(NQES_EINVAL, "W is NULL");
    if (C < 1 || m < 1 || B < 0) return c->fail(NQES_EINVAL, "invalid C=%d m=%d B=%d", C, m, B);
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    const int N = c->N, K = c->K;
    const long long ncol = (2LL * N + 2LL * N * N) * K * K;
    DevBuf dcor, dpart, dsum;
    HIPCHK(c, dcor.ensure((size_t)C * ncol * 16), "alloc correlations");
    const int nch = choose_chunks(C, ncol, 64);
    HIPCHK(c, dpart.ensure((size_t)nch * ncol * 16), "alloc correlations");
    HIPCHK(c, dsum.ensure((size_t)ncol * 16), "alloc correlations");
    HIPCHK(c, c->dspins.ensure(16), "alloc");
    HIPCHK(c, hipMemcpyAsync(c->dW.p, W, 16 * c->P, hipMemcpyHostToDevice, c->stream), "copy W");
    HIPCHK(c, nq_launch_tables64(N, c->M, K, c->L, c->dW.p, c->dchT.p, c->dshT.p, c->deaP.p, c->deaN.p, c->stream),
           "tables kernel");
    Fp64SampleArgs a{};
    memset(&a, 0, sizeof(a));
    a.N = N;
    a.M = c->M;
    a.NW = c->NW;
    a.L = c->L;
    a.P = c->P;
    a.C = C;
    a.m = m;
    a.B = B;
    a.seed = seed;
    a.h = c->h;
    a.model = c->model;
    a.delta = c->model == NQES_MODEL_XXZ ? -1.0 : 1.0;
    a.J = static_cast<const double *>(c->dJ.p);
    a.W = c->dW.p;
    a.chT = c->dchT.p;
    a.shT = c->dshT.p;
    a.eaP = c->deaP.p;
    a.eaN = c->deaN.p;
    a.cor = dcor.p;
    HIPCHK(c, nq_launch_corr64(K, a, c->stream), "correlation kernel");
    HIPCHK(c, nq_launch_colsum64(dcor.p, C, ncol, nch, dpart.p, dsum.p, c->stream), "colsum kernel");
    std::vector<cd> sum(ncol);
    HIPCHK(c, hipMemcpyAsync(sum.data(), dsum.p, (size_t)ncol * 16, hipMemcpyDeviceToHost, c->stream), "copy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "correlation sync");
    const double Sd = (double)C * (double)m;
    const long long kk2 = (long long)K * K;
    double *outs[4] = {x, z, xx, zz};
    const long long off[4] = {0, (long long)N, 2LL * N, 2LL * N + (long long)N * N};
    const long long cnt[4] = {(long long)N, (long long)N, (long long)N * N, (long long)N * N};
    for (int t = 0; t < 4; t++) {
        if (!outs[t]) continue;
        for (long long q = 0; q < cnt[t] * kk2; q++) {
            cd v = nq64::divr(sum[off[t] * kk2 + q], Sd);
            outs[t][2 * q] = v.re;
            outs[t][2 * q + 1] = v.im;
        }
    }
    return NQES_OK;
}

int nqes_sr_step(nqes_ctx *c, double *W, uint64_t seed, int C, int m, int B, double lr, double reglam, double rtol,
                 uint64_t maxit, double *emat, nqes_stats *stats)
{
    if (!c) return NQES_EINVAL;
    if (!W || !emat) return c->fail(NQES_EINVAL, "NULL W or emat");
    auto t0 = std::chrono::steady_clock::now();
    int st = nqes_sample(c, W, seed, C, m, B, reglam);
    if (st != NQES_OK) return st;
    double t_sample = event_ms(c->ev[0], c->ev[1]);
    double t_mom = event_ms(c->ev[1], c->ev[2]);
    auto t1 = std::chrono::steady_clock::now();
    std::vector<cd> x;
    nqmq::Result r;
    r.flag = 0;
    r.iters = 0;
    if (!c->nonfinite) {
        st = solve_impl(c, rtol, maxit, x, r);
        if (st != NQES_OK) return st;
    }
    auto t2 = std::chrono::steady_clock::now();
    memcpy(emat, c->emat.data(), 16 * c->K * c->K);
    if (c->nonfinite) {
        if (stats) stats->nonfinite = 1;
        if (c->theta_overflow)
            return c->fail(NQES_ENONFINITE,
                           "non-finite local energy or gradient: a hidden-unit angle |theta| > 80 exceeds the mixed "
                           "path's FP32 range (the FP64 path has no such limit)");
        return c->fail(NQES_ENONFINITE, "non-finite local energy or gradient");
    }
    // weight update (rbm.cc:699-717)
    cd *Wc = reinterpret_cast<cd *>(W);
    for (long long q = 0; q < c->P; q++) Wc[q] = nq64::sub(Wc[q], nq64::scale(lr, x[q]));
    if (stats) {
        std::vector<unsigned long long> acc(C), ties(C);
        HIPCHK(c, copy_sync(acc.data(), c->dacc.p, (size_t)C * 8, hipMemcpyDeviceToHost), "copy accepts");
        HIPCHK(c, copy_sync(ties.data(), c->dties.p, (size_t)C * 8, hipMemcpyDeviceToHost), "copy tie counts");
        uint64_t a = 0, nt = 0;
        for (auto v : acc) a += v;
        for (auto v : ties) nt += v;
        stats->ulp_ties = nt;
        stats->minres_flag = r.flag;
        stats->minres_iters = r.iters;
        stats->accepts = a;
        stats->proposals = (uint64_t)C * (uint64_t)(B + m) * (uint64_t)c->K * (uint64_t)c->N;
        stats->sv_calls = c->sv_calls;
        stats->t_sample_ms = t_sample;
        stats->t_moments_ms = t_mom;
        stats->t_solve_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        stats->t_sv_ms = c->sv_ms;
        stats->t_total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
        stats->nonfinite = 0;
    }
    return NQES_OK;
}

} // extern "C"
