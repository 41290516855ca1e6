// sampler_mx.hip -- Metropolis sweeps of the mixed path (rbm.cc:530-552).
//
// G lanes per chain, 64/G chains per wavefront, all chains of a wavefront flipping
// the same (replica k, spin i) in lockstep.  Lane g of a chain owns hidden units
// j = g + G q (q < JQ) of every replica pair: their tanh(theta_{k,kk,j}) live in
// registers, so a flip touches memory only for the table row of spin i.  Tables are
// padded to Mp = G JQ hidden units (cosh = 1, sinh = w = b = 0 beyond M, which makes
// the padded units neutral: tanh 0, ratio factor 1) and interleaved so that one
// 8-byte (16-byte complex) load gives the cosh/sinh pair of a unit.
//
// Per flip of spin i of replica k:
//   den_j = cosh(2w) - t_j sinh(+-2w),   t'_j = (t_j cosh(2w) - sinh(+-2w)) / den_j,
//   Phi'_{k,kk} = Phi_{k,kk} exp(-+2a_i) prod_j den_j   (FP32 within 8 lanes, then FP64),
//   det Phi' = sum_kk Phi'_{k,kk} cof_{k,kk}(Phi)       (Laplace along row k),
// and the move is taken if u |det Phi|^2 <= |det Phi'|^2.  The cofactors of row k do
// not depend on row k, so they are formed once per replica and sweep.  After every
// sweep theta, tanh and Phi are recomputed from the spins (drift bounded to one sweep).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "cplx64.hpp"
#include "kernels.hpp"
#include "mixed_math.hpp"

using namespace nqm;
using nq64::cd;

namespace {

template <typename T> struct ChSh;
template <> struct __align__(8) ChSh<float> {
    float ch, sh;
};
template <> struct __align__(16) ChSh<cf> {
    cf ch, sh;
};

__device__ inline float coshT(float x) { return coshf(x); }
__device__ inline cf coshT(cf z)
{
    float sy, cy;
    __sincosf(z.im, &sy, &cy);
    float e = __expf(z.re), ie = __expf(-z.re);
    return mkf(0.5f * (e + ie) * cy, 0.5f * (e - ie) * sy);
}
__device__ inline float reT(float x) { return x; }
__device__ inline float reT(cf z) { return z.re; }
__device__ inline double expD(double x) { return nq64::exp_(x); }
__device__ inline cd expD(cd z) { return nq64::cexp_(z); }

// padded table set: D eaP[K N], eaN[K N], aD[K N] (section rounded up to 16 bytes);
// ChSh cs[K N Mp]; T w[K N Mp]; T b[K Mp].  Mp is a multiple of 32, so the total is a
// multiple of 16 bytes.
__host__ __device__ inline size_t pad_tab_dbytes(bool cplx, int N, int K)
{
    return (3 * (size_t)K * N * (cplx ? 16 : 8) + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t pad_tab_bytes(bool cplx, int N, int K, int Mp)
{
    const size_t ts = cplx ? 8 : 4;
    return pad_tab_dbytes(cplx, N, K) + (size_t)K * N * Mp * 3 * ts + (size_t)K * Mp * ts;
}

#if !defined(NQES_SWEEP_K) || NQES_SWEEP_K == 1
template <bool CPLX>
__global__ void k_pad_tables_mx(int N, int M, int K, int Mp, const void *chT_, const void *shT_, const void *wT_,
                                const void *bT_, const void *eaP_, const void *eaN_, const void *aD_, void *out)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const T *chT = (const T *)chT_, *shT = (const T *)shT_, *wT = (const T *)wT_, *bT = (const T *)bT_;
    const long long KN = (long long)K * N, KNM = KN * Mp;
    D *d = (D *)out;
    ChSh<T> *cs = (ChSh<T> *)((unsigned char *)out + pad_tab_dbytes(CPLX, N, K));
    T *w = (T *)(cs + KNM);
    T *b = w + KNM;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < KNM; q += stride) {
        const long long row = q / Mp;
        const int j = (int)(q % Mp);
        ChSh<T> v;
        v.ch = oneT<T>();
        v.sh = zeroT<T>();
        T wv = zeroT<T>();
        if (j < M) {
            v.ch = chT[row * M + j];
            v.sh = shT[row * M + j];
            wv = wT[row * M + j];
        }
        cs[q] = v;
        w[q] = wv;
        if (q < KN) {
            d[q] = ((const D *)eaP_)[q];
            d[KN + q] = ((const D *)eaN_)[q];
            d[2 * KN + q] = ((const D *)aD_)[q];
        }
        if (q < (long long)K * Mp) {
            const int kk = (int)(q / Mp), jj = (int)(q % Mp);
            b[q] = jj < M ? bT[(long long)kk * M + jj] : zeroT<T>();
        }
    }
}
#endif

// cofactors of row k (a runtime index) of the K x K matrix phi (row-major, LDS):
// det = sum_c phi[k][c] cof[c].  The K - 1 other rows are read by computed LDS
// addresses, so the minors are formed with static register indices.
template <int K, typename D>
__device__ inline void cof_row(const D *phi, int k, D *cof)
{
    if constexpr (K == 1) {
        cof[0] = oneD<D>();
    } else {
        D rows[(K - 1) * K];
#pragma unroll
        for (int r = 0; r < K - 1; r++) {
            const int src = r < k ? r : r + 1;
#pragma unroll
            for (int cc = 0; cc < K; cc++) rows[r * K + cc] = phi[src * K + cc];
        }
#pragma unroll
        for (int cc = 0; cc < K; cc++) {
            D mn[(K - 1) * (K - 1)];
            int o = 0;
#pragma unroll
            for (int r = 0; r < K - 1; r++)
#pragma unroll
                for (int c2 = 0; c2 < K; c2++) {
                    if (c2 == cc) continue;
                    mn[o++] = rows[r * K + c2];
                }
            const D v = det_m<K - 1, D>(mn);
            cof[cc] = ((k + cc) & 1) ? negD(v) : v;
        }
    }
}

// Per-chain state in LDS between replica blocks: tanh(theta) [k][kk][q][g] (lane g
// touches only its own entries) and Phi [k][kk] (written by lane 0 of the group, read
// by all of it; one wavefront, so program order orders the accesses).
template <bool CPLX, int K, int G, int JQ>
__host__ __device__ constexpr size_t chain_lds_bytes(int NW)
{
    return (size_t)K * K * JQ * G * (CPLX ? 8 : 4) + (size_t)K * K * (CPLX ? 16 : 8) + (size_t)K * NW * 4;
}

// Theta -> tanh and Phi of the chain held by a group of G lanes, from Theta0 of chain
// c (chain start) or from the spins sp; results to the chain's LDS state.
template <bool CPLX, int K, int G, int JQ>
__device__ __attribute__((always_inline)) inline void sweep_refresh(const MixedSampleArgs &a, const typename Ty<CPLX>::D *aD,
                                     const typename Ty<CPLX>::T *wp, const typename Ty<CPLX>::T *bp,
                                     const uint32_t *sp, typename Ty<CPLX>::T *tls, typename Ty<CPLX>::D *phis,
                                     int g, long long c, bool from_theta0, int &overflow)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    constexpr int Mp = G * JQ;
    const int N = a.N, M = a.M, NW = a.NW;
    T th[K][K][JQ];
    if (from_theta0) {
        const T *th0 = (const T *)a.theta0;
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int kk = 0; kk < K; kk++)
#pragma unroll
                for (int q = 0; q < JQ; q++) {
                    const int j = g + G * q;
                    th[k][kk][q] = j < M ? th0[((c * K + k) * K + kk) * M + j] : zeroT<T>();
                }
    } else {
#pragma unroll
        for (int kk = 0; kk < K; kk++)
#pragma unroll
            for (int q = 0; q < JQ; q++) {
                const T b = bp[kk * Mp + g + G * q];
#pragma unroll
                for (int k = 0; k < K; k++) th[k][kk][q] = b;
            }
        for (int w0 = 0; w0 < NW; w0++) {
            uint32_t bits[K];
#pragma unroll
            for (int k = 0; k < K; k++) bits[k] = sp[k * NW + w0];
            const int iend = (N - 32 * w0 < 32) ? N - 32 * w0 : 32;
#pragma unroll 4
            for (int ii = 0; ii < iend; ii++) {
                const int i = 32 * w0 + ii;
                T wv[K][JQ];
#pragma unroll
                for (int kk = 0; kk < K; kk++)
#pragma unroll
                    for (int q = 0; q < JQ; q++) wv[kk][q] = wp[((long long)kk * N + i) * Mp + g + G * q];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const bool dn = (bits[k] >> ii) & 1u;
#pragma unroll
                    for (int kk = 0; kk < K; kk++)
#pragma unroll
                        for (int q = 0; q < JQ; q++)
                            th[k][kk][q] = dn ? subT(th[k][kk][q], wv[kk][q]) : addT(th[k][kk][q], wv[kk][q]);
                }
            }
        }
    }
    // angles to the chain's LDS state, then one (replica, RBM) pair at a time: tanh in
    // place, Phi_{k,kk} = exp(sum_i +-a_{kk,i}) prod_j cosh(theta_j)
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int kk = 0; kk < K; kk++)
#pragma unroll
            for (int q = 0; q < JQ; q++) tls[((k * K + kk) * JQ + q) * G + g] = th[k][kk][q];
#pragma unroll 1
    for (int pr = 0; pr < K * K; pr++) {
        const int k = pr / K, kk = pr % K;
        D part = oneD<D>();
#pragma unroll
        for (int q = 0; q < JQ; q++) {
            T *tp = tls + (pr * JQ + q) * G + g;
            const T t = *tp;
            if (fabsf(reT(t)) > 80.0f) overflow = 1;
            *tp = tanhT(t);
            part = mulD(part, toD(coshT(t)));
        }
        // sum_i +-a_{kk,i}: lane g takes i = g, g + G, ..., then a sum over the group
        D A = zeroD<D>();
        for (int i = g; i < N; i += G) {
            const D ai = aD[kk * N + i];
            A = ((sp[k * NW + (i >> 5)] >> (i & 31)) & 1u) ? subD(A, ai) : addD(A, ai);
        }
        A = group_sum<G>(A);
        const D ph = mulD(expD(A), group_prod<G>(part));
        if (g == 0) phis[pr] = ph;
    }
    __builtin_amdgcn_wave_barrier();
}

// Metropolis chains (see the file comment).  Tables in LDS when TAB, else read from
// the padded global copy (L2-resident).  Replica k's angles and Phi row are held in
// registers for its block of N flips.  After each sampled sweep the spins are stored
// (sample s = c m + t); the local energies and O rows are computed afterwards by
// k_measure_mx, one lane group per sample.
template <bool CPLX, int K, int G, int JQ, bool TAB>
__global__ __launch_bounds__(256) void k_sample_mx(MixedSampleArgs a, const void *tabp)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    constexpr int CPW = 64 / G;
    constexpr int Mp = G * JQ;
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, NW = a.NW;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane & (G - 1), grp = lane / G;
    const int cpb = (blockDim.x >> 6) * CPW;
    const int cib = wv * CPW + grp;
    const int c_raw = blockIdx.x * cpb + cib;
    const bool live = c_raw < a.C;
    const int c = live ? c_raw : a.C - 1;
    const long long KN = (long long)K * N;

    // LDS: [padded tables] (TAB) [chain states]
    const D *dtab;
    unsigned char *p = smem;
    if constexpr (TAB) {
        const size_t nb = pad_tab_bytes(CPLX, N, K, Mp);
        const uint4 *src = (const uint4 *)tabp;
        uint4 *dst = (uint4 *)smem;
        for (size_t q = tid; q < nb / 16; q += blockDim.x) dst[q] = src[q];
        dtab = (const D *)smem;
        p = smem + nb;
    } else {
        dtab = (const D *)tabp;
    }
    const D *eaP = dtab, *eaN = dtab + KN, *aD = dtab + 2 * KN;
    const ChSh<T> *cs = (const ChSh<T> *)((const unsigned char *)dtab + pad_tab_dbytes(CPLX, N, K));
    const T *wp = (const T *)(cs + KN * Mp);
    const T *bp = wp + KN * Mp;
    unsigned char *st = p + (size_t)cib * chain_lds_bytes<CPLX, K, G, JQ>(NW);
    D *phis = (D *)st;
    T *tls = (T *)(phis + K * K);
    uint32_t *sp = (uint32_t *)(tls + K * K * JQ * G);

    if (g == 0)
        for (int w = 0; w < K * NW; w++) sp[w] = a.spins0[(long long)c * K * NW + w];
    nq64::pcg rng;
    rng.lo = a.rng0[2 * c];
    rng.hi = a.rng0[2 * c + 1];
    __syncthreads();

    int overflow = 0;
    // Draws in batches of G: lane g of the group holds draw g of the batch, computed
    // with one jump of g + 1 steps from the batch base (the same stream as G
    // sequential steps); proposal t of the batch reads lane t's draw.
    nq64::pcg jmp_a, jmp_c;
    nq64::pcg_jump((unsigned)g + 1u, jmp_a, jmp_c);
    double ub = 0.0;
    int tdraw = G;
    unsigned long long accepts = 0;
    const int nsweeps = a.B + a.m;
    for (int sw = 0; sw < nsweeps; sw++) {
        // angles and Phi from Theta0 (first sweep) or recomputed from the spins; a
        // chain's LDS state is touched by its own wavefront only, so no block barrier
        sweep_refresh<CPLX, K, G, JQ>(a, aD, wp, bp, sp, tls, phis, g, c, sw == 0, overflow);
        D det_phi;
        {
            D ph[K * K];
#pragma unroll
            for (int q = 0; q < K * K; q++) ph[q] = phis[q];
            det_phi = det_m<K, D>(ph);
        }
#pragma unroll 1
        for (int k = 0; k < K; k++) {
            D cof[K], prc[K];
            T tk[K][JQ];
            {
                cof_row<K, D>(phis, k, cof);
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    prc[kk] = phis[k * K + kk];
#pragma unroll
                    for (int q = 0; q < JQ; q++) tk[kk][q] = tls[((k * K + kk) * JQ + q) * G + g];
                }
            }
            // table row of the next flip, loaded one flip ahead from global memory
            // (L2 latency); from LDS it is loaded at the top of its own flip
            ChSh<T> nx[TAB ? 1 : K][TAB ? 1 : JQ];
            if constexpr (!TAB) {
#pragma unroll
                for (int kk = 0; kk < K; kk++)
#pragma unroll
                    for (int q = 0; q < JQ; q++) nx[kk][q] = cs[((long long)kk * N) * Mp + g + G * q];
            }
            uint32_t cur = sp[k * NW];
            for (int i = 0; i < N; i++) {
                ChSh<T> cv[K][JQ];
                if constexpr (TAB) {
#pragma unroll
                    for (int kk = 0; kk < K; kk++)
#pragma unroll
                        for (int q = 0; q < JQ; q++) cv[kk][q] = cs[((long long)kk * N + i) * Mp + g + G * q];
                } else {
                    const int i1 = (i + 1 < N) ? i + 1 : i;
#pragma unroll
                    for (int kk = 0; kk < K; kk++)
#pragma unroll
                        for (int q = 0; q < JQ; q++) {
                            cv[kk][q] = nx[kk][q];
                            nx[kk][q] = cs[((long long)kk * N + i1) * Mp + g + G * q];
                        }
                }
                if (i > 0 && (i & 31) == 0) cur = sp[k * NW + (i >> 5)];
                const bool si = (cur >> (i & 31)) & 1u;
                const D *ea = si ? eaP : eaN;
                T tn[K][JQ];
                T pf[K];
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    pf[kk] = oneT<T>();
#pragma unroll
                    for (int q = 0; q < JQ; q++) {
                        const T chv = cv[kk][q].ch;
                        const T shv = si ? negT(cv[kk][q].sh) : cv[kk][q].sh;
                        const T th = tk[kk][q];
                        const T den = subT(chv, mulT(th, shv));
                        tn[kk][q] = divT(subT(mulT(th, chv), shv), den);
                        pf[kk] = mulT(pf[kk], den);
                    }
                }
                D prow[K];
                D det_p = zeroD<D>();
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    const D Tp = group_prod_tf<G, T, D>(pf[kk]);
                    prow[kk] = mulD(mulD(prc[kk], ea[kk * N + i]), Tp);
                    det_p = addD(det_p, mulD(prow[kk], cof[kk]));
                }
                if (tdraw == G) {  // next batch of G draws (wave-uniform: chains step in lockstep)
                    const nq64::pcg sg = nq64::pcg_affine(rng, jmp_a, jmp_c);
                    ub = nq64::canonical(nq64::pcg_out(sg));
                    rng.lo = __shfl(sg.lo, G - 1, G);
                    rng.hi = __shfl(sg.hi, G - 1, G);
                    tdraw = 0;
                }
                // draw tdraw of each group's batch: scalar reads of lane grp G + tdraw
                double u = 0.0;
#pragma unroll
                for (int q = 0; q < CPW; q++) {
                    const uint64_t ubits = __builtin_bit_cast(uint64_t, ub);
                    const uint32_t lo32 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)ubits, q * G + tdraw);
                    const uint32_t hi32 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(ubits >> 32), q * G + tdraw);
                    if (grp == q) u = __builtin_bit_cast(double, (uint64_t)lo32 | ((uint64_t)hi32 << 32));
                }
                tdraw++;
                // U <= min(1, |det'|^2 / |det|^2) without the division (rbm.cc:542-543)
                const bool acc = isZeroD(det_phi) || u * normD(det_phi) <= normD(det_p);
                if (acc) {
                    det_phi = det_p;
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        prc[kk] = prow[kk];
#pragma unroll
                        for (int q = 0; q < JQ; q++) tk[kk][q] = tn[kk][q];
                    }
                    cur ^= 1u << (i & 31);
                    accepts++;
                }
                if (((i & 31) == 31 || i == N - 1) && g == 0) sp[k * NW + (i >> 5)] = cur;
            }
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                if (g == 0) phis[k * K + kk] = prc[kk];
#pragma unroll
                for (int q = 0; q < JQ; q++) tls[((k * K + kk) * JQ + q) * G + g] = tk[kk][q];
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (sw >= a.B && live) {
            const long long s = (long long)c * a.m + (sw - a.B);
            for (int w = g; w < K * NW; w += G) a.spins[s * K * NW + w] = sp[w];
        }
    }
    if (live && g == 0) a.accepts[c] = accepts;
    if (live && overflow) atomicOr(a.flags, 1u);
}

template <bool CPLX, int K, int G, int JQ>
hipError_t launch_jq(const MixedSampleArgs &a, const void *tabp, bool tab, int wpb, hipStream_t st)
{
    const int cpb = wpb * (64 / G);
    size_t lds = (size_t)cpb * chain_lds_bytes<CPLX, K, G, JQ>(a.NW);
    if (tab) lds += pad_tab_bytes(CPLX, a.N, K, G * JQ);
    const int blocks = (a.C + cpb - 1) / cpb;
    const void *fn = tab ? (const void *)k_sample_mx<CPLX, K, G, JQ, true> : (const void *)k_sample_mx<CPLX, K, G, JQ, false>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (tab) hipLaunchKernelGGL((k_sample_mx<CPLX, K, G, JQ, true>), dim3(blocks), dim3(64 * wpb), lds, st, a, tabp);
    else hipLaunchKernelGGL((k_sample_mx<CPLX, K, G, JQ, false>), dim3(blocks), dim3(64 * wpb), lds, st, a, tabp);
    return hipGetLastError();
}

template <bool CPLX, int K, int G>
hipError_t launch_g(const MixedSampleArgs &a, const void *tabp, bool tab, int JQ, int wpb, hipStream_t st)
{
    switch (JQ) {
    case 1: return launch_jq<CPLX, K, G, 1>(a, tabp, tab, wpb, st);
    case 2: return launch_jq<CPLX, K, G, 2>(a, tabp, tab, wpb, st);
    case 4: return launch_jq<CPLX, K, G, 4>(a, tabp, tab, wpb, st);
    case 8: return launch_jq<CPLX, K, G, 8>(a, tabp, tab, wpb, st);
    default: return hipErrorInvalidValue;
    }
}

template <bool CPLX, int K>
hipError_t launch_k(const MixedSampleArgs &a, const void *tabp, const SweepShape &sh, hipStream_t st)
{
    switch (sh.G) {
    case 32: return launch_g<CPLX, K, 32>(a, tabp, sh.tab, sh.JQ, sh.wpb, st);
    case 64: return launch_g<CPLX, K, 64>(a, tabp, sh.tab, sh.JQ, sh.wpb, st);
    default: return hipErrorInvalidValue;
    }
}

} // namespace

// This file is compiled once per replica count (NQES_SWEEP_K = 1 .. 4) so that the
// instantiations build in parallel; the K = 1 unit also holds the shape and dispatch.
#ifndef NQES_SWEEP_K
#define NQES_SWEEP_K 1
#endif

template <>
hipError_t nq_sweep_k<NQES_SWEEP_K>(bool cplx, const MixedSampleArgs &a, const void *tabp, const SweepShape &sh,
                                    hipStream_t st)
{
    return cplx ? launch_k<true, NQES_SWEEP_K>(a, tabp, sh, st) : launch_k<false, NQES_SWEEP_K>(a, tabp, sh, st);
}

#if NQES_SWEEP_K == 1

// G = 32 lanes per chain (measured faster than 64 at K = 2 and 4), 64 when M > 256;
// NQES_G = 32 / 64 overrides.  JQ = ceil(M / G) rounded up to
// 1, 2, 4 or 8.  Tables in LDS when they fit next to the spins.
static int round_jq(int M, int G)
{
    const int jpl = (M + G - 1) / G;
    int JQ = 1;
    while (JQ < jpl) JQ *= 2;
    return JQ;
}

SweepShape nq_sweep_shape(bool cplx, int N, int M, int K, size_t lds_max)
{
    SweepShape sh{};
    int G = round_jq(M, 32) > 8 ? 64 : 32;
    const char *e = getenv("NQES_G");
    if (e && (atoi(e) == 32 || atoi(e) == 64) && round_jq(M, atoi(e)) <= 8) G = atoi(e);
    const int JQ = round_jq(M, G);
    if (JQ > 8) return sh;
    int wpb = 4;
    const char *ew = getenv("NQES_WPB");
    if (ew && *ew && atoi(ew) >= 1 && atoi(ew) <= 4) wpb = atoi(ew);
    const int NW = (N + 31) / 32;
    const size_t per_chain =
        (size_t)K * K * JQ * G * (cplx ? 8 : 4) + (size_t)K * K * (cplx ? 16 : 8) + (size_t)K * NW * 4;
    auto states = [&](int w) { return (size_t)w * (64 / G) * per_chain; };
    const size_t tb = pad_tab_bytes(cplx, N, K, G * JQ);
    if (lds_max > 0) {
        while (wpb > 1 && states(wpb) > lds_max) wpb--;
        if (states(wpb) > lds_max) return sh;
    }
    sh.ok = 1;
    sh.G = G;
    sh.JQ = JQ;
    sh.wpb = wpb;
    const char *et = getenv("NQES_TABLDS");
    sh.tab = (tb + states(wpb) <= lds_max) && !(et && *et && atoi(et) == 0);
    sh.tab_bytes = pad_tab_bytes(cplx, N, K, G * JQ);
    return sh;
}

hipError_t nq_launch_sweep_mx(bool cplx, int K, const MixedSampleArgs &a, const SweepShape &sh, void *tabp,
                              hipStream_t st)
{
    if (!sh.ok) return hipErrorInvalidValue;
    const int Mp = sh.G * sh.JQ;
    const long long n = (long long)K * a.N * Mp;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (cplx)
        hipLaunchKernelGGL(k_pad_tables_mx<true>, dim3(blocks), dim3(256), 0, st, a.N, a.M, K, Mp, a.chT, a.shT, a.wT,
                           a.bT, a.eaP, a.eaN, a.aD, tabp);
    else
        hipLaunchKernelGGL(k_pad_tables_mx<false>, dim3(blocks), dim3(256), 0, st, a.N, a.M, K, Mp, a.chT, a.shT,
                           a.wT, a.bT, a.eaP, a.eaN, a.aD, tabp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    switch (K) {
    case 1: return nq_sweep_k<1>(cplx, a, tabp, sh, st);
    case 2: return nq_sweep_k<2>(cplx, a, tabp, sh, st);
    case 3: return nq_sweep_k<3>(cplx, a, tabp, sh, st);
    case 4: return nq_sweep_k<4>(cplx, a, tabp, sh, st);
    default: return hipErrorInvalidValue;
    }
}
#endif
