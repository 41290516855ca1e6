// kernels_mixed.hip -- mixed-precision device path of one SR step.
//
// Per-hidden-unit work in FP32 (tanh theta, the cosh/sinh update factors, O),
// per-chain K x K algebra, amplitude ratios and every sum in FP64.  CPLX = false
// is the real specialisation used when all weights are real.
//
// Kernels
//   k_tables_mx      FP32 cosh(2w), sinh(2w) and w, FP64 exp(+-2a), a, b (once per step)
//   k_chain_init     per chain: pcg64(seed + c), K*N Bernoulli draws (rbm.cc:99-119)
//   k_theta_gemm     Theta0[C*K, K*M] = Sigma[C*K, N] W[N, K*M] + b on FP32 MFMA
//                    (v_mfma_f32_16x16x4f32), the angles at chain start (rbm.cc:135-145)
//   (k_sample_mx and k_measure_mx, the Metropolis sweeps and the local energies with
//   the O rows, are in sampler_mx.hip)
//   k_moments_mx     column sums of O, |O|^2, E conj(O) in FP64 with Neumaier compensation
//   k_sv_rows_mx     S*v in one streamed pass over O (rbm.cc:661-684): WPR wavefronts
//                    share a row, w_s = (O_s - o)^T v, then acc += conj(O_s) w_s with
//                    FP64 accumulation in registers; next row prefetched, nontemporal loads
//   k_sv_rowdot_mx / k_sv_colacc_mx   two-pass S*v when v does not fit in LDS
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cplx64.hpp"
#include "kernels.hpp"
#include "mixed_math.hpp"

hipError_t nq_allow_dynamic_lds(const void *fn, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;  // allowed without the attribute
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, size_t> allowed;  // (device, kernel) -> bytes
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    auto it = allowed.find({dev, fn});
    if (it == allowed.end()) {
        NqUnsafeHold hold;
        int lds_max = 0;
        e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        if (e != hipSuccess) return e;
        hipFuncAttributes fa{};
        e = hipFuncGetAttributes(&fa, fn);
        if (e != hipSuccess) return e;
        const size_t most = (size_t)lds_max > fa.sharedSizeBytes ? (size_t)lds_max - fa.sharedSizeBytes : 0;
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)most);
        if (e != hipSuccess) return e;
        it = allowed.emplace(std::make_pair(dev, fn), most).first;
    }
    return bytes <= it->second ? hipSuccess : hipErrorInvalidValue;
}

using namespace nqm;
using nq64::cd;

namespace {

__device__ inline double fromW(cd w, double) { return w.re; }
__device__ inline cd fromW(cd w, cd) { return w; }

// ---------------------------------------------------------------- tables
template <bool CPLX>
__global__ void k_tables_mx(int N, int M, int K, int L, const cd *__restrict__ W, void *chT_, void *shT_, void *wT_,
                            void *bT_, void *eaP_, void *eaN_, void *aD_)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    T *chT = (T *)chT_, *shT = (T *)shT_, *wT = (T *)wT_, *bT = (T *)bT_;
    D *eaP = (D *)eaP_, *eaN = (D *)eaN_, *aD = (D *)aD_;
    long long n = (long long)K * N * M;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        int j = (int)(idx % M);
        long long r = idx / M;
        int i = (int)(r % N);
        int kk = (int)(r / N);
        cd w = W[(long long)kk * L + N + M + (long long)i * M + j];
        cd w2 = nq64::scale(2.0, w);
        chT[idx] = toT(fromW(nq64::ccosh_(w2), D()));
        shT[idx] = toT(fromW(nq64::csinh_(w2), D()));
        wT[idx] = toT(fromW(w, D()));
        if (i == 0) bT[kk * M + j] = toT(fromW(W[(long long)kk * L + N + j], D()));
        if (j == 0) {
            cd a = W[(long long)kk * L + i];
            cd a2 = nq64::scale(2.0, a);
            eaP[kk * N + i] = fromW(nq64::cexp_(a2), D());
            eaN[kk * N + i] = fromW(nq64::cexp_(nq64::neg(a2)), D());
            aD[kk * N + i] = fromW(a, D());
        }
    }
}

// ---------------------------------------------------------------- chain start
__global__ void k_chain_init(int N, int K, int NW, int C, uint64_t seed, uint32_t *__restrict__ spins0,
                             uint64_t *__restrict__ rng0)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    nq64::pcg rng = nq64::pcg_seed(seed + (uint64_t)c);
    for (int w = 0; w < K * NW; w++) {
        uint32_t word = 0;
        int i0 = (w % NW) * 32;
        for (int b = 0; b < 32 && i0 + b < N; b++)
            if (nq64::canonical(nq64::pcg_next(rng)) < 0.5) word |= 1u << b;
        spins0[(long long)c * K * NW + w] = word;
    }
    rng0[2 * c] = rng.lo;
    rng0[2 * c + 1] = rng.hi;
}

// Theta0 = Sigma W + b with FP32 MFMA 16x16x4: one wavefront per 16x16 output tile.
// Rows r = c*K + k (chain c, replica k), columns col over K*M*R (R = 2 for complex:
// re/im interleaved, so the output is the [c][k][kk][j] complex layout).
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_theta_gemm(int N, int M, int K, int NW, int R, long long rows,
                                                   const uint32_t *__restrict__ spins0,
                                                   const float *__restrict__ wT, const float *__restrict__ bT,
                                                   float *__restrict__ theta0)
{
    const int lane = threadIdx.x;
    const long long ncol = (long long)K * M * R;
    const long long r0 = (long long)blockIdx.x * 16;
    const long long c0 = (long long)blockIdx.y * 16;
    const long long arow = r0 + (lane & 15);
    const long long bcol = c0 + (lane & 15);
    const int kq = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    // B column bcol -> (kk, j, part)
    int bkk = 0, bj = 0, bpart = 0;
    bool bok = bcol < ncol;
    if (bok) {
        long long e = bcol / R;
        bpart = (int)(bcol % R);
        bkk = (int)(e / M);
        bj = (int)(e % M);
    }
    const bool aok = arow < rows;
    const uint32_t *sp = spins0 + (aok ? arow : 0) * NW;  // row = c*K + k -> spins0[c][k][word]
    for (int i0 = 0; i0 < N; i0 += 4) {
        int i = i0 + kq;
        float av = 0.f, bv = 0.f;
        if (aok && i < N) av = ((sp[i >> 5] >> (i & 31)) & 1u) ? -1.f : 1.f;
        if (bok && i < N) bv = wT[(((long long)bkk * N + i) * M + bj) * R + bpart];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
    // lane holds D[4*(lane/16) + q][lane%16]
    const long long ocol = c0 + (lane & 15);
    if (ocol < ncol) {
        long long e = ocol / R;
        int part = (int)(ocol % R);
        int kk = (int)(e / M), j = (int)(e % M);
        float bias = bT[((long long)kk * M + j) * R + part];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            long long orow = r0 + 4 * (lane >> 4) + q;
            if (orow < rows) theta0[orow * ncol + ocol] = acc[q] + bias;
        }
    }
}

// Theta0 = Sigma W + b on BF16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
// W (FP32) is split into three bf16 terms w1 + w2 + w3: w1 = bf16(W), w2 = bf16(W - w1),
// w3 = bf16(W - w1 - w2); each residual holds at most the 16 and 8 significant bits left,
// so it is exact in FP32 and w3 is exact in bf16: W = w1 + w2 + w3 exactly (normal
// numbers).  The spins are +-1 (exact in bf16), so the MFMAs sum exact products in FP32,
// as the FP32 form does, in another order.  Block: 4 wavefronts over a 64-row x 64-column
// output tile; the three terms of the block's 64 columns in LDS as bf16, [term][col][k]
// with k contiguous (N padded to a multiple of 32 with zeros, rows padded by 8 bf16 so
// that the 16 lanes of a group read different banks).  Wavefront w: rows 16 w .. 16 w +
// 15 of the tile, its 4 column tiles.  Operands: lane l, element e = K index 8 (l/16) + e
// of a 32-wide step (the same bijection for A and B); result register e of lane l = row
// 4 (l/16) + e, column l % 16.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ inline uint16_t bf16_rne(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__global__ __launch_bounds__(256) void k_theta_bf16(int N, int M, int K, int NW, int R, long long rows,
                                                   const uint32_t *__restrict__ spins0,
                                                   const float *__restrict__ wT, const float *__restrict__ bT,
                                                   float *__restrict__ theta0)
{
    extern __shared__ __align__(16) uint16_t wl[];  // [3][64][NPP]
    const int NP = (N + 31) & ~31, NPP = NP + 8;
    const long long ncol = (long long)K * M * R;
    const long long c0 = (long long)blockIdx.y * 64;
    for (int q = threadIdx.x; q < 64 * NP; q += blockDim.x) {
        const int i = q >> 6, cl = q & 63;  // consecutive threads: consecutive columns of W row i
        const long long col = c0 + cl;
        float w = 0.f;
        if (col < ncol && i < N) {
            const long long e = col / R;
            const int part = (int)(col % R), kk = (int)(e / M), j = (int)(e % M);
            w = wT[(((long long)kk * N + i) * M + j) * R + part];
        }
        const uint16_t h1 = bf16_rne(w);
        const float r1 = w - bf16_to_f(h1);
        const uint16_t h2 = bf16_rne(r1);
        const uint16_t h3 = bf16_rne(r1 - bf16_to_f(h2));
        wl[(0 * 64 + cl) * NPP + i] = h1;
        wl[(1 * 64 + cl) * NPP + i] = h2;
        wl[(2 * 64 + cl) * NPP + i] = h3;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, lo = lane & 15;
    // row tiles rt0, rt0 + gridDim.x, ... of 64 rows share the staged W terms
    for (long long rtile = blockIdx.x; rtile * 64 < rows; rtile += gridDim.x) {
        const long long r0 = rtile * 64;
        const long long arow = r0 + 16 * wv + lo;
        const bool aok = arow < rows;
        const uint32_t *sp = spins0 + (aok ? arow : 0) * NW;
        v4f acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ct++) acc[ct] = v4f{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < NP; k0 += 32) {
            const int i0 = k0 + 8 * g;  // this lane's 8 spins
            const uint32_t word = (aok && i0 < N) ? sp[i0 >> 5] : 0u;
            union {
                uint16_t h[8];
                v8bf v;
            } a;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int i = i0 + e;
                a.h[e] = (aok && i < N) ? (((word >> (i & 31)) & 1u) ? (uint16_t)0xBF80 : (uint16_t)0x3F80) : (uint16_t)0;
            }
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const v8bf b = *(const v8bf *)(wl + (t * 64 + 16 * ct + lo) * NPP + i0);
                    acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b, acc[ct], 0, 0, 0);
                }
        }
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            const long long ocol = c0 + 16 * ct + lo;
            if (ocol >= ncol) continue;
            const long long e = ocol / R;
            const int part = (int)(ocol % R), kk = (int)(e / M), j = (int)(e % M);
            const float bias = bT[((long long)kk * M + j) * R + part];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long long orow = r0 + 16 * wv + 4 * g + q;
                if (orow < rows) theta0[orow * ncol + ocol] = acc[ct][q] + bias;
            }
        }
    }
}

// ---------------------------------------------------------------- moments
__device__ inline void neum_add(double &s, double &comp, double x)
{
    double t = s + x;
    if (fabs(s) >= fabs(x)) comp += (s - t) + x;
    else comp += (x - t) + s;
    s = t;
}

template <bool CPLX>
__global__ void k_moments_mx(const void *O_, const cd *__restrict__ E, long long S, long long P, long long Pld,
                             int nchunk, cd *__restrict__ po, cd *__restrict__ ps, cd *__restrict__ pf)
{
    using T = typename Ty<CPLX>::T;
    const T *O = (const T *)O_;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    long long s0 = S * blockIdx.y / nchunk, s1 = S * (blockIdx.y + 1) / nchunk;
    double o_r = 0, o_rc = 0, o_i = 0, o_ic = 0, s_r = 0, s_rc = 0, f_r = 0, f_rc = 0, f_i = 0, f_ic = 0;
    for (long long s = s0; s < s1; s++) {
        cd v;
        if constexpr (CPLX) v = toD(O[s * Pld + p]);
        else v = nq64::mk(toD(O[s * Pld + p]), 0.0);
        const cd e = E[s];
        neum_add(o_r, o_rc, v.re);
        neum_add(o_i, o_ic, v.im);
        neum_add(s_r, s_rc, v.re * v.re + v.im * v.im);
        // E conj(v)
        neum_add(f_r, f_rc, e.re * v.re + e.im * v.im);
        neum_add(f_i, f_ic, e.im * v.re - e.re * v.im);
    }
    long long o = (long long)blockIdx.y * P + p;
    po[o] = nq64::mk(o_r + o_rc, o_i + o_ic);
    ps[o] = nq64::mk(s_r + s_rc, 0.0);
    pf[o] = nq64::mk(f_r + f_rc, f_i + f_ic);
}

// ---------------------------------------------------------------- S*v (one pass)
// A 16-byte vector of an O row holds 4 real or 2 complex columns.
template <bool CPLX> struct Vec;
template <> struct Vec<false> {
    static constexpr int W = 4;
};
template <> struct Vec<true> {
    static constexpr int W = 2;
};

template <bool CPLX>
__device__ inline void vec_to_D(const float4 &x, typename Ty<CPLX>::D *o)
{
    if constexpr (CPLX) {
        o[0] = nq64::mk(x.x, x.y);
        o[1] = nq64::mk(x.z, x.w);
    } else {
        o[0] = x.x;
        o[1] = x.y;
        o[2] = x.z;
        o[3] = x.w;
    }
}

typedef float f4v __attribute__((ext_vector_type(4)));

// streaming load that bypasses cache retention (O is read once per S*v)
__device__ __forceinline__ float4 ld_stream(const float4 *p)
{
    f4v r = __builtin_nontemporal_load((const f4v *)p);
    return make_float4(r.x, r.y, r.z, r.w);
}

// Row-group S*v.  A block holds NP row groups of WPR wavefronts.  Lane l of wave h
// of a group owns the 16-byte vectors u = l + 64 (h + WPR q), q < QV, of the group's
// rows; group g of block b takes rows b NP + g + k (grid NP).  Per row: partial dot
// with v (FP64, staged once per block in LDS), 64-lane butterfly, the WPR partials
// summed in wave order through LDS (one barrier per row when WPR > 1; the slots
// alternate between two buffers so one barrier suffices), then acc += conj(O_s) w_s
// in FP64 registers.  The next row's loads are issued before the current row's
// math.  Every wave of a block runs the same number of iterations (rows past S
// contribute w = 0), so the barriers match.  Output: one partial row of Pld values
// per block (groups summed in order) and the block's sum of w.
template <bool CPLX, int QV, int WPR, int NP>
__global__ __launch_bounds__(64 * WPR * NP) void k_sv_rows_mx(const float4 *__restrict__ O, const void *v_,
                                                               const void *ov_, long long S, long long nvec,
                                                               void *part_, void *wsum_)
{
    using D = typename Ty<CPLX>::D;
    constexpr int W = Vec<CPLX>::W;
    extern __shared__ __align__(16) unsigned char smem_sv[];
    __shared__ D red[2][NP][WPR];
    __shared__ D wred[NP];
    D *vl = (D *)smem_sv;
    const D *v = (const D *)v_;
    const D ov = *(const D *)ov_;
    const int t = threadIdx.x, lane = t & 63, h = (t >> 6) % WPR, g = t / (64 * WPR);
    for (long long q = t; q < nvec * W; q += blockDim.x) vl[q] = v[q];
    __syncthreads();
    D acc[QV][W];
#pragma unroll
    for (int q = 0; q < QV; q++)
#pragma unroll
        for (int e = 0; e < W; e++) acc[q][e] = zeroD<D>();
    D wacc = zeroD<D>();
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    const long long ns = (long long)gridDim.x * NP;
    const long long s0 = (long long)blockIdx.x * NP;
    long long s = s0 + g;
    float4 x[QV];
#pragma unroll
    for (int q = 0; q < QV; q++) {
        long long u = lane + 64LL * (h + WPR * q);
        x[q] = (s < S && u < nvec) ? ld_stream(O + s * nvec + u) : z4;
    }
    int par = 0;
    for (long long sb = s0; sb < S; sb += ns, s += ns) {
        float4 y[QV];
        const long long sn = s + ns;
#pragma unroll
        for (int q = 0; q < QV; q++) {
            long long u = lane + 64LL * (h + WPR * q);
            y[q] = (sn < S && u < nvec) ? ld_stream(O + sn * nvec + u) : z4;
        }
        D dot = zeroD<D>();
#pragma unroll
        for (int q = 0; q < QV; q++) {
            long long u = lane + 64LL * (h + WPR * q);
            if (u < nvec) {
                D od[W];
                vec_to_D<CPLX>(x[q], od);
#pragma unroll
                for (int e = 0; e < W; e++) dot = addD(dot, mulD(od[e], vl[u * W + e]));
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) dot = addD(dot, shflxD(dot, off));
        if constexpr (WPR > 1) {
            if (lane == 0) red[par][g][h] = dot;
            __syncthreads();
            dot = red[par][g][0];
#pragma unroll
            for (int k = 1; k < WPR; k++) dot = addD(dot, red[par][g][k]);
            par ^= 1;
        }
        const D w = (s < S) ? subD(dot, ov) : zeroD<D>();
        wacc = addD(wacc, w);
#pragma unroll
        for (int q = 0; q < QV; q++) {
            // re-convert from the FP32 registers instead of keeping FP64 copies alive
            float4 xx = x[q];
            asm volatile("" : "+v"(xx.x), "+v"(xx.y), "+v"(xx.z), "+v"(xx.w));
            D od[W];
            vec_to_D<CPLX>(xx, od);
#pragma unroll
            for (int e = 0; e < W; e++) acc[q][e] = addD(acc[q][e], mulD(w, conjD(od[e])));
        }
#pragma unroll
        for (int q = 0; q < QV; q++) x[q] = y[q];
    }
    // sum the groups' accumulators in group order through LDS (reusing the v area)
    for (int k = 0; k < NP; k++) {
        __syncthreads();
        if (g == k) {
#pragma unroll
            for (int q = 0; q < QV; q++) {
                long long u = lane + 64LL * (h + WPR * q);
                if (u < nvec)
#pragma unroll
                    for (int e = 0; e < W; e++) vl[u * W + e] = (k == 0) ? acc[q][e] : addD(vl[u * W + e], acc[q][e]);
            }
        }
    }
    if (h == 0 && lane == 0) wred[g] = wacc;
    __syncthreads();
    D *part = (D *)part_ + (long long)blockIdx.x * nvec * W;
    for (long long q = t; q < nvec * W; q += blockDim.x) part[q] = vl[q];
    if (t == 0) {
        D ws = wred[0];
        for (int k = 1; k < NP; k++) ws = addD(ws, wred[k]);
        ((D *)wsum_)[blockIdx.x] = ws;
    }
}

// two-pass fallback: w_s = (O_s - o)^T v (one wave per row), then column sums
template <bool CPLX>
__global__ __launch_bounds__(256) void k_sv_rowdot_mx(const void *O_, const void *v_, const void *ov_, long long S,
                                                      long long P, long long Pld, void *w_)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const T *O = (const T *)O_;
    const D *v = (const D *)v_;
    const D ov = *(const D *)ov_;
    D *w = (D *)w_;
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long s = wave; s < S; s += nwaves) {
        const T *row = O + s * Pld;
        D acc = zeroD<D>();
        for (long long p = lane; p < P; p += 64) acc = addD(acc, mulD(toD(row[p]), v[p]));
        acc = group_sum<64>(acc);
        if (lane == 0) w[s] = subD(acc, ov);
    }
}

template <bool CPLX>
__global__ void k_sv_colacc_mx(const void *O_, const void *w_, long long S, long long P, long long Pld, int nchunk,
                               void *part_, void *wpart_)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const T *O = (const T *)O_;
    const D *w = (const D *)w_;
    D *part = (D *)part_;
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long s0 = S * blockIdx.y / nchunk, s1 = S * (blockIdx.y + 1) / nchunk;
    if (p < P) {
        D acc = zeroD<D>();
        for (long long s = s0; s < s1; s++) acc = addD(acc, mulD(w[s], conjD(toD(O[s * Pld + p]))));
        part[(long long)blockIdx.y * P + p] = acc;
    }
    if (p == 0) {
        D ws = zeroD<D>();
        for (long long s = s0; s < s1; s++) ws = addD(ws, w[s]);
        ((D *)wpart_)[blockIdx.y] = ws;
    }
}

// acc + part[c0] + ... + part[c1 - 1] (rows of part, stride apart), added in order; the
// loads go eight ahead of the adds (a dependent load per add would wait out its latency)
template <typename D>
__device__ inline D chunk_run(D acc, const D *__restrict__ part, long long stride, int c0, int c1)
{
    int ch = c0;
    for (; ch + 8 <= c1; ch += 8) {
        D x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = part[(long long)(ch + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) acc = addD(acc, x[u]);
    }
    for (; ch < c1; ch++) acc = addD(acc, part[(long long)ch * stride]);
    return acc;
}

// out[col] = sum over chunks of part[chunk * stride + col], col < ncol, chunks in order;
// blockIdx.z: the set (part + z pss, out + z oss)
template <typename D>
__global__ void k_chunksum_mx(const D *__restrict__ part, long long ncol, long long stride, int nchunk,
                              D *__restrict__ out, long long pss, long long oss)
{
    long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    part += blockIdx.z * pss;
    out[blockIdx.z * oss + col] = chunk_run(part[col], part + col, stride, 1, nchunk);
}

// partial sums over groups of CG consecutive chunks: grid (col tiles, groups, sets)
template <typename D>
__global__ void k_chunkgroup_mx(const D *__restrict__ part, long long ncol, long long stride, int nchunk, int cg,
                                D *__restrict__ gout, long long pss)
{
    long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    int c0 = blockIdx.y * cg, c1 = c0 + cg;
    if (c1 > nchunk) c1 = nchunk;
    gout[((long long)blockIdx.z * gridDim.y + blockIdx.y) * ncol + col] =
        chunk_run(zeroD<D>(), part + blockIdx.z * pss + col, stride, c0, c1);
}

// out[col] = sum_chunks part[...], deterministic, parallel over column tiles and
// chunk groups; scratch holds ngroups * ncol values per set.  nsets independent sets of
// partial rows (part + z pss) reduce into out + z oss in one pair of launches.
template <typename D>
void chunk_reduce(const D *part, long long ncol, long long stride, int nchunk, D *scratch, D *out, hipStream_t st,
                  int nsets = 1, long long pss = 0, long long oss = 0)
{
    const int cg = 32;
    int ngroups = (nchunk + cg - 1) / cg;
    if (ngroups <= 1) {
        hipLaunchKernelGGL(k_chunksum_mx<D>, dim3((unsigned)((ncol + 63) / 64), 1, (unsigned)nsets), dim3(64), 0, st,
                           part, ncol, stride, nchunk, out, pss, oss);
        return;
    }
    hipLaunchKernelGGL(k_chunkgroup_mx<D>, dim3((unsigned)((ncol + 63) / 64), (unsigned)ngroups, (unsigned)nsets),
                       dim3(64), 0, st, part, ncol, stride, nchunk, cg, scratch, pss);
    hipLaunchKernelGGL(k_chunksum_mx<D>, dim3((unsigned)((ncol + 63) / 64), 1, (unsigned)nsets), dim3(64), 0, st,
                       (const D *)scratch, ncol, ncol, ngroups, out, (long long)ngroups * ncol, oss);
}

} // namespace

hipError_t nq_chunk_reduce(bool cplx, const void *part, long long ncol, long long stride, int nchunk,
                           void *scratch, void *out, hipStream_t st)
{
    if (cplx) chunk_reduce<cd>((const cd *)part, ncol, stride, nchunk, (cd *)scratch, (cd *)out, st);
    else chunk_reduce<double>((const double *)part, ncol, stride, nchunk, (double *)scratch, (double *)out, st);
    return hipGetLastError();
}

hipError_t nq_chunk_reduce_sets(const double *part, long long ncol, long long stride, int nchunk, double *scratch,
                                double *out, int nsets, long long pss, long long oss, hipStream_t st)
{
    chunk_reduce<double>(part, ncol, stride, nchunk, scratch, out, st, nsets, pss, oss);
    return hipGetLastError();
}

hipError_t nq_launch_tables_mx(bool cplx, int N, int M, int K, int L, const void *W, void *chT, void *shT, void *wT,
                               void *bT, void *eaP, void *eaN, void *aD, hipStream_t st)
{
    long long n = (long long)K * N * M;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    if (cplx)
        hipLaunchKernelGGL(k_tables_mx<true>, dim3(blocks), dim3(256), 0, st, N, M, K, L, (const cd *)W, chT, shT, wT,
                           bT, eaP, eaN, aD);
    else
        hipLaunchKernelGGL(k_tables_mx<false>, dim3(blocks), dim3(256), 0, st, N, M, K, L, (const cd *)W, chT, shT,
                           wT, bT, eaP, eaN, aD);
    return hipGetLastError();
}

hipError_t nq_launch_chain_init(int N, int K, int NW, int C, uint64_t seed, uint32_t *spins0, uint64_t *rng0,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_chain_init, dim3((C + 127) / 128), dim3(128), 0, st, N, K, NW, C, seed, spins0, rng0);
    return hipGetLastError();
}

hipError_t nq_launch_theta_gemm(bool cplx, int N, int M, int K, int NW, int C, const uint32_t *spins0,
                                const void *wT, const void *bT, void *theta0, hipStream_t st)
{
    int R = cplx ? 2 : 1;
    long long rows = (long long)C * K;
    long long ncol = (long long)K * M * R;
    const char *e = getenv("NQES_THETA");  // "fp32": the FP32-MFMA form
    if (!(e && strcmp(e, "fp32") == 0)) {
        const int NPP = ((N + 31) & ~31) + 8;
        // the W terms (101 KB at N = 256) or the output tile, whichever is larger
        const size_t lds = (size_t)3 * 64 * NPP * 2;  // 101 KB at N = 256
        if (lds > 65536) {
            hipError_t ea = nq_allow_dynamic_lds((const void *)k_theta_bf16, lds);
            if (ea != hipSuccess) return ea;
        }
        // about 2048 blocks: a block loops over row tiles so that large sizes stage each
        // W slice once per several tiles
        const long long cb = (ncol + 63) / 64, rtiles = (rows + 63) / 64;
        long long rb = 2048 / cb;
        if (rb < 1) rb = 1;
        if (rb > rtiles) rb = rtiles;
        dim3 grid((unsigned)rb, (unsigned)cb);
        hipLaunchKernelGGL(k_theta_bf16, grid, dim3(256), lds, st, N, M, K, NW, R, rows, spins0, (const float *)wT,
                           (const float *)bT, (float *)theta0);
        return hipGetLastError();
    }
    dim3 grid((unsigned)((rows + 15) / 16), (unsigned)((ncol + 15) / 16));
    hipLaunchKernelGGL(k_theta_gemm, grid, dim3(64), 0, st, N, M, K, NW, R, rows, spins0, (const float *)wT,
                       (const float *)bT, (float *)theta0);
    return hipGetLastError();
}

hipError_t nq_launch_moments_mx(bool cplx, const void *O, const void *E, long long S, long long P, long long Pld,
                                int nchunk, void *po, void *ps, void *pf, void *scratch, void *out_o, void *out_s,
                                void *out_f, hipStream_t st)
{
    dim3 g((unsigned)((P + 255) / 256), (unsigned)nchunk);
    if (cplx)
        hipLaunchKernelGGL(k_moments_mx<true>, g, dim3(256), 0, st, O, (const cd *)E, S, P, Pld, nchunk, (cd *)po,
                           (cd *)ps, (cd *)pf);
    else
        hipLaunchKernelGGL(k_moments_mx<false>, g, dim3(256), 0, st, O, (const cd *)E, S, P, Pld, nchunk, (cd *)po,
                           (cd *)ps, (cd *)pf);
    // the partial buffers' first chunk rows double as scratch once reduced in order
    chunk_reduce<cd>((const cd *)po, P, P, nchunk, (cd *)scratch, (cd *)out_o, st);
    chunk_reduce<cd>((const cd *)ps, P, P, nchunk, (cd *)scratch, (cd *)out_s, st);
    chunk_reduce<cd>((const cd *)pf, P, P, nchunk, (cd *)scratch, (cd *)out_f, st);
    return hipGetLastError();
}

// Shape of the row-group S*v for rows of nvec 16-byte vectors: the fewest waves per
// row (1, 2, 4, 8) that keep a lane's share at <= 9 vectors, NP groups so that a block
// has 4 waves (8 when WPR = 8), and 2048 waves in the grid (two blocks per CU at the
// kernel's two waves per SIMD).  qv = 0: the row is too long (two-pass form).
// NQES_SV_WPR sets the smallest WPR tried and NQES_SV_BLOCKS the block count.
SvShape nq_sv_shape(bool cplx, long long nvec)
{
    SvShape sh{0, 0, 0, 0, 0};
    size_t lds = (size_t)nvec * 32;  // v in FP64: 4 doubles or 2 complex per vector
    if (lds > 160 * 1024) return sh;
    const char *ew = getenv("NQES_SV_WPR");
    const int wmin = (ew && *ew) ? atoi(ew) : 1;
    for (int wpr = 1; wpr <= 8; wpr *= 2) {
        long long q = (nvec + 64LL * wpr - 1) / (64LL * wpr);
        if (q <= 9 && wpr >= wmin) {
            sh.wpr = wpr;
            sh.qv = (int)q;
            break;
        }
    }
    if (!sh.qv) return sh;
    if (sh.wpr > 1 && sh.qv < 5) sh.qv = 5;  // instantiated shapes only; idle slots are masked
    sh.np = sh.wpr >= 4 ? 1 : 4 / sh.wpr;
    sh.blocks = 2048 / (sh.wpr * sh.np);
    const char *e = getenv("NQES_SV_BLOCKS");
    if (e && *e && atoi(e) > 0) sh.blocks = atoi(e);
    sh.lds = lds;
    (void)cplx;
    return sh;
}

template <bool CPLX, int QV, int WPR, int NP>
hipError_t launch_rows(const void *O, const void *v, const void *ov, long long S, long long nvec, int nblk,
                       void *part, void *wpart, size_t lds, hipStream_t st)
{
    if (lds > 65536) {
        hipError_t e = nq_allow_dynamic_lds((const void *)k_sv_rows_mx<CPLX, QV, WPR, NP>, lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sv_rows_mx<CPLX, QV, WPR, NP>), dim3(nblk), dim3(64 * WPR * NP), lds, st,
                       (const float4 *)O, v, ov, S, nvec, part, wpart);
    return hipGetLastError();
}

template <bool CPLX>
hipError_t launch_rows_shape(const SvShape &sh, const void *O, const void *v, const void *ov, long long S,
                             long long nvec, int nblk, void *part, void *wpart, hipStream_t st)
{
#define NQ_R(q, w, n)                                                                                             \
    if (sh.qv == q && sh.wpr == w && sh.np == n)                                                                 \
        return launch_rows<CPLX, q, w, n>(O, v, ov, S, nvec, nblk, part, wpart, sh.lds, st);
    NQ_R(1, 1, 4) NQ_R(2, 1, 4) NQ_R(3, 1, 4) NQ_R(4, 1, 4) NQ_R(5, 1, 4) NQ_R(6, 1, 4) NQ_R(7, 1, 4)
    NQ_R(8, 1, 4) NQ_R(9, 1, 4)
    NQ_R(5, 2, 2) NQ_R(6, 2, 2) NQ_R(7, 2, 2) NQ_R(8, 2, 2) NQ_R(9, 2, 2)
    NQ_R(5, 4, 1) NQ_R(6, 4, 1) NQ_R(7, 4, 1) NQ_R(8, 4, 1) NQ_R(9, 4, 1)
    NQ_R(5, 8, 1) NQ_R(6, 8, 1) NQ_R(7, 8, 1) NQ_R(8, 8, 1) NQ_R(9, 8, 1)
#undef NQ_R
    return hipErrorInvalidValue;
}

// S*v raw sums: out[p] = sum_s conj(O_sp) w_s (P entries of D), out[P] = sum_s w_s,
// w_s = O_s^T v - ov.  v, ov, out are D (double or complex double).  nblk partial
// rows: the row-group kernel's block count (nq_sv_shape) or, for rows too long for
// it, the chunk count of the two-pass form.  scratch: >= (nblk/32 + 2) * (Pld + 1) D.
hipError_t nq_launch_sv_mx(bool cplx, const void *O, const void *v, const void *ov, long long S, long long P,
                           long long Pld, int nblk, void *part, void *wpart, void *wbuf, void *scratch, void *out,
                           hipStream_t st)
{
    const int W = cplx ? 2 : 4;
    const long long nvec = Pld / W;
    hipError_t e = hipSuccess;
    auto reduce = [&](long long ncol, long long stride) {
        if (cplx) {
            chunk_reduce<cd>((const cd *)part, ncol, stride, nblk, (cd *)scratch, (cd *)out, st);
            chunk_reduce<cd>((const cd *)wpart, 1, 1, nblk, (cd *)scratch + (long long)(nblk / 32 + 1) * ncol,
                             (cd *)out + P, st);
        } else {
            chunk_reduce<double>((const double *)part, ncol, stride, nblk, (double *)scratch, (double *)out, st);
            chunk_reduce<double>((const double *)wpart, 1, 1, nblk,
                                 (double *)scratch + (long long)(nblk / 32 + 1) * ncol, (double *)out + P, st);
        }
        return hipGetLastError();
    };
    const SvShape sh = nq_sv_shape(cplx, nvec);
    if (sh.qv) {
        e = cplx ? launch_rows_shape<true>(sh, O, v, ov, S, nvec, nblk, part, wpart, st)
                 : launch_rows_shape<false>(sh, O, v, ov, S, nvec, nblk, part, wpart, st);
        if (e != hipSuccess) return e;
        return reduce(P, Pld);
    }
    long long blocks = (S + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    dim3 g((unsigned)((P + 255) / 256), (unsigned)nblk);
    if (cplx) {
        hipLaunchKernelGGL(k_sv_rowdot_mx<true>, dim3((unsigned)blocks), dim3(256), 0, st, O, v, ov, S, P, Pld, wbuf);
        hipLaunchKernelGGL(k_sv_colacc_mx<true>, g, dim3(256), 0, st, O, wbuf, S, P, Pld, nblk, part, wpart);
    } else {
        hipLaunchKernelGGL(k_sv_rowdot_mx<false>, dim3((unsigned)blocks), dim3(256), 0, st, O, v, ov, S, P, Pld,
                           wbuf);
        hipLaunchKernelGGL(k_sv_colacc_mx<false>, g, dim3(256), 0, st, O, wbuf, S, P, Pld, nblk, part, wpart);
    }
    return reduce(P, P);
}
