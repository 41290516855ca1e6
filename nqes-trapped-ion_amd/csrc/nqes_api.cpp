// nqes_api.cpp -- host side of libnqes: the C ABI of include/nqes.h.
//
// One SR step (reference: monte_carlo, code/rbm.cc:564-720):
//   1. weights -> device, per-step tables (k_tables64)
//   2. sampler kernel: one wavefront per chain, writes O (S x P) to HBM, H_loc, E_loc
//   3. deterministic column reductions -> sums of O, |O|^2, E O*, H_loc, E_loc
//   4. host normalisation (rbm.cc:641-650)
//   5. MINRES-QLP on the host (FP64 scalars, length-P vectors), whose operator
//      is the two streamed S*v passes over O on the device (rbm.cc:661-684)
//   6. w -= lr * x (rbm.cc:699-717)
// Host arithmetic is compiled with -ffp-contract=off as the device code is.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nqes.h"
#include "cplx64.hpp"
#include "kernels.hpp"
#include "minresqlp.hpp"

using nq64::cd;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n)
    {
        if (n <= bytes && p) return hipSuccess;
        release();
        if (n == 0) n = 16;
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

} // namespace

struct nqes_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;

    // problem
    bool has_problem = false;
    int N = 0, M = 0, K = 0, NW = 0, L = 0, model = 0, prec = 0;
    long long P = 0;
    double h = 0;
    DevBuf dJ, dW, dchT, dshT, deaP, deaN;

    // sample state
    bool sampled = false;
    int C = 0, m = 0, B = 0;
    long long S = 0;
    DevBuf dO, dHs, dEs, dspins, dacc;
    DevBuf dmpart_o, dmpart_s, dmpart_f, dmo, dms, dmf, dhpart, dhsum, depart, desum;
    DevBuf dv, dw, dsvpart, dsvout;
    int mchunks = 1, svchunks = 1;
    std::vector<cd> o, skk, f, emat;
    cd eloc{0, 0};
    std::vector<cd> svraw;
    uint64_t sv_calls = 0;
    double sv_ms = 0;
    int nonfinite = 0;

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

extern "C" {

const char *nqes_version(void) { return "nqes-mi355x 0.1 (gfx950)"; }

nqes_ctx *nqes_create(int device)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return nullptr;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    nqes_ctx *c = new nqes_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return nullptr;
        }
    return c;
}

void nqes_destroy(nqes_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *nqes_last_error(const nqes_ctx *c) { return c ? c->err.c_str() : "null context"; }

int64_t nqes_required_bytes(int N, int M, int K, int C, int m, int precision)
{
    if (N <= 0 || M <= 0 || K <= 0 || C <= 0 || m <= 0) return -1;
    long long P = (long long)K * (N + M + (long long)N * M);
    long long S = (long long)C * m;
    long long NW = (N + 31) / 32;
    long long b = 0;
    b += S * P * 16;                  // O
    b += S * (long long)K * K * 16;   // H_loc
    b += S * 16 * 2;                  // E_loc, w
    b += S * K * NW * 4;              // spins
    b += (long long)K * N * M * 32;   // tables
    b += 3 * 4096LL * P * 16;         // chunk partials (upper bound)
    (void)precision;
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
    if (model != NQES_MODEL_TRAPPEDION) return c->fail(NQES_EINVAL, "unsupported model %d", model);
    if (precision != NQES_PREC_FP64) return c->fail(NQES_EINVAL, "unsupported precision %d", precision);
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
    size_t lds = (size_t)16 * K * K * M + 4 * (size_t)K * ((N + 31) / 32);
    if (lds > prop.sharedMemPerBlock)
        return c->fail(NQES_EINVAL, "K*K*M=%d hidden-unit angles (%zu B) exceed the %zu B of LDS per block", K * K * M,
                       lds, (size_t)prop.sharedMemPerBlock);
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
    HIPCHK(c, hipMemcpy(c->dJ.p, J, sizeof(double) * N * N, hipMemcpyHostToDevice), "copy J");
    HIPCHK(c, c->dW.ensure(16 * c->P), "alloc W");
    HIPCHK(c, c->dchT.ensure((size_t)16 * K * N * M), "alloc tables");
    HIPCHK(c, c->dshT.ensure((size_t)16 * K * N * M), "alloc tables");
    HIPCHK(c, c->deaP.ensure((size_t)16 * K * N), "alloc tables");
    HIPCHK(c, c->deaN.ensure((size_t)16 * K * N), "alloc tables");
    c->has_problem = true;
    c->sampled = false;
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
    size_t freeb = 0, totalb = 0;
    HIPCHK(c, hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo");
    long long need = (long long)(S * P * 16) - (long long)c->dO.bytes;
    if (need > 0 && (size_t)need + (size_t)(1ull << 30) > freeb)
        return c->fail(NQES_ENOMEM, "O needs %.2f GB, %.2f GB free", S * P * 16 / 1e9, freeb / 1e9);
    HIPCHK(c, c->dO.ensure((size_t)(S * P * 16)), "alloc O");
    HIPCHK(c, c->dHs.ensure((size_t)(S * K * K * 16)), "alloc H");
    HIPCHK(c, c->dEs.ensure((size_t)(S * 16)), "alloc E");
    HIPCHK(c, c->dspins.ensure((size_t)(S * K * c->NW * 4)), "alloc spins");
    HIPCHK(c, c->dacc.ensure((size_t)C * 8), "alloc accepts");
    c->mchunks = choose_chunks(S, P, 256);
    HIPCHK(c, c->dmpart_o.ensure((size_t)c->mchunks * P * 16), "alloc partials");
    HIPCHK(c, c->dmpart_s.ensure((size_t)c->mchunks * P * 16), "alloc partials");
    HIPCHK(c, c->dmpart_f.ensure((size_t)c->mchunks * P * 16), "alloc partials");
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
    HIPCHK(c, nq_launch_tables64(c->N, c->M, K, c->L, c->dW.p, c->dchT.p, c->dshT.p, c->deaP.p, c->deaN.p, c->stream),
           "tables kernel");
    Fp64SampleArgs a;
    a.N = c->N;
    a.M = c->M;
    a.NW = c->NW;
    a.L = c->L;
    a.P = P;
    a.C = C;
    a.m = m;
    a.B = B;
    a.seed = seed;
    a.h = c->h;
    a.J = static_cast<const double *>(c->dJ.p);
    a.W = c->dW.p;
    a.chT = c->dchT.p;
    a.shT = c->dshT.p;
    a.eaP = c->deaP.p;
    a.eaN = c->deaN.p;
    a.O = c->dO.p;
    a.Hs = c->dHs.p;
    a.Es = c->dEs.p;
    a.spins = static_cast<uint32_t *>(c->dspins.p);
    a.accepts = static_cast<unsigned long long *>(c->dacc.p);
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream), "event");
    HIPCHK(c, nq_launch_sample64(K, a, c->stream), "sample kernel");
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream), "event");
    HIPCHK(c, nq_launch_moments64(c->dO.p, c->dEs.p, S, P, c->mchunks, c->dmpart_o.p, c->dmpart_s.p, c->dmpart_f.p,
                                  c->dmo.p, c->dms.p, c->dmf.p, c->stream),
           "moments kernel");
    HIPCHK(c, nq_launch_colsum64(c->dHs.p, S, K * K, hch, c->dhpart.p, c->dhsum.p, c->stream), "colsum kernel");
    HIPCHK(c, nq_launch_colsum64(c->dEs.p, S, 1, ech, c->depart.p, c->desum.p, c->stream), "colsum kernel");
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");

    std::vector<cd> so(P), ss(P), sf(P), sh(K * K), se(1);
    HIPCHK(c, hipMemcpyAsync(so.data(), c->dmo.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(ss.data(), c->dms.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(sf.data(), c->dmf.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(sh.data(), c->dhsum.p, 16 * K * K, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipMemcpyAsync(se.data(), c->desum.p, 16, hipMemcpyDeviceToHost, c->stream), "copy moments");
    HIPCHK(c, hipStreamSynchronize(c->stream), "sample sync");

    // normalisation (rbm.cc:641-650)
    using namespace nq64;
    const double Sd = (double)S;
    c->eloc = divr(se[0], Sd);
    c->o.resize(P);
    c->skk.resize(P);
    c->f.resize(P);
    int nonfinite = !(std::isfinite(c->eloc.re) && std::isfinite(c->eloc.im));
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
    c->C = C;
    c->m = m;
    c->B = B;
    c->S = S;
    c->svchunks = choose_chunks(S, P, 256);
    HIPCHK(c, c->dv.ensure((size_t)P * 16), "alloc v");
    HIPCHK(c, c->dw.ensure((size_t)S * 16), "alloc w");
    HIPCHK(c, c->dsvpart.ensure((size_t)c->svchunks * P * 16), "alloc S*v partials");
    HIPCHK(c, c->dsvout.ensure((size_t)P * 16), "alloc S*v out");
    c->svraw.resize(P);
    c->sv_calls = 0;
    c->sv_ms = 0;
    c->sampled = true;
    return NQES_OK;
}

int nqes_get_samples(nqes_ctx *c, uint32_t *spins, double *H, uint64_t *accepts)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    if (spins)
        HIPCHK(c, hipMemcpy(spins, c->dspins.p, (size_t)c->S * c->K * c->NW * 4, hipMemcpyDeviceToHost), "copy spins");
    if (H) HIPCHK(c, hipMemcpy(H, c->dHs.p, (size_t)c->S * c->K * c->K * 16, hipMemcpyDeviceToHost), "copy H");
    if (accepts) HIPCHK(c, hipMemcpy(accepts, c->dacc.p, (size_t)c->C * 8, hipMemcpyDeviceToHost), "copy accepts");
    return NQES_OK;
}

int nqes_get_O(nqes_ctx *c, int64_t row0, int64_t nrows, double *out)
{
    if (!c) return NQES_EINVAL;
    if (!c->sampled) return c->fail(NQES_ESTATE, "no samples: call nqes_sample first");
    if (!out || row0 < 0 || nrows < 0 || row0 + nrows > c->S) return c->fail(NQES_EINVAL, "row range out of bounds");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    HIPCHK(c,
           hipMemcpy(out, static_cast<char *>(c->dO.p) + (size_t)row0 * c->P * 16, (size_t)nrows * c->P * 16,
                     hipMemcpyDeviceToHost),
           "copy O");
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

// S v = (1/S) sum_s (O_s^T v) conj(O_s) + skk .* v - (o^T v) conj(o)   (rbm.cc:661-684)
static int apply_S(nqes_ctx *c, const cd *v, cd *out)
{
    const long long P = c->P;
    HIPCHK(c, hipMemcpyAsync(c->dv.p, v, 16 * P, hipMemcpyHostToDevice, c->stream), "copy v");
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream), "event");
    HIPCHK(c, nq_launch_sv64(c->dO.p, c->dv.p, c->S, P, c->svchunks, c->dw.p, c->dsvpart.p, c->dsvout.p, c->stream),
           "S*v kernels");
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream), "event");
    HIPCHK(c, hipMemcpyAsync(c->svraw.data(), c->dsvout.p, 16 * P, hipMemcpyDeviceToHost, c->stream), "copy S*v");
    HIPCHK(c, hipStreamSynchronize(c->stream), "S*v sync");
    c->sv_ms += event_ms(c->ev[2], c->ev[3]);
    c->sv_calls++;
    using namespace nq64;
    cd ov = zero();
    for (long long q = 0; q < P; q++) ov = add(ov, mul(c->o[q], v[q]));
    const double Sd = (double)c->S;
    for (long long q = 0; q < P; q++) {
        cd r = divr(c->svraw[q], Sd);
        out[q] = add(r, sub(mul(c->skk[q], v[q]), mul(ov, conj(c->o[q]))));
    }
    return NQES_OK;
}

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
    int status = NQES_OK;
    bool aborted = false;
    auto op = [&](const std::vector<cd> &v, std::vector<cd> &out) -> bool {
        status = apply_S(c, v.data(), out.data());
        return status == NQES_OK;
    };
    res = nqmq::solve(op, c->f, x, rtol, maxit, aborted);
    if (aborted) return status;
    return NQES_OK;
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
        return c->fail(NQES_ENONFINITE, "non-finite local energy or gradient");
    }
    // weight update (rbm.cc:699-717)
    cd *Wc = reinterpret_cast<cd *>(W);
    for (long long q = 0; q < c->P; q++) Wc[q] = nq64::sub(Wc[q], nq64::scale(lr, x[q]));
    if (stats) {
        std::vector<unsigned long long> acc(C);
        HIPCHK(c, hipMemcpy(acc.data(), c->dacc.p, (size_t)C * 8, hipMemcpyDeviceToHost), "copy accepts");
        uint64_t a = 0;
        for (auto v : acc) a += v;
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
