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

#include <type_traits>

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

// Padded table set (byte offsets from its start, every section 16-byte aligned):
//   D   eaP[K N], eaN[K N], aD[K N]
//   cs  ChSh [K N][Mp + 1]   cosh/sinh pairs; the extra entry per row (1, 0) makes
//                            the row stride odd, so lanes reading one hidden unit of
//                            different rows hit different LDS banks
//   w   T [K N][Mp],  b  T [K][Mp]
// The measurement kernel reads a second set with the same layout except that its cs
// section is transposed, ChSh [K][Mp][N] (spin fastest): its lanes run over spins, so
// the pairs that one load instruction reads are adjacent (one cache line per 8 or 16
// lanes instead of one per lane when the tables stay in L2).
struct PadLayout {
    size_t cs, w, b, total;
};
__host__ __device__ inline size_t al16(size_t x) { return (x + 15) & ~(size_t)15; }
__host__ __device__ inline PadLayout pad_layout(bool cplx, int N, int K, int Mp)
{
    const size_t ts = cplx ? 8 : 4, ds = cplx ? 16 : 8;
    PadLayout L;
    L.cs = al16(3 * (size_t)K * N * ds);
    L.w = L.cs + al16((size_t)K * N * (Mp + 1) * 2 * ts);
    L.b = L.w + (size_t)K * N * Mp * ts;
    L.total = al16(L.b + (size_t)K * Mp * ts);
    return L;
}
__host__ __device__ inline size_t pad_tab_bytes(bool cplx, int N, int K, int Mp) { return pad_layout(cplx, N, K, Mp).total; }

#if !defined(NQES_SWEEP_K) || NQES_SWEEP_K == 1
template <bool CPLX>
__global__ void k_pad_tables_mx(int N, int M, int K, int Mp, const void *chT_, const void *shT_, const void *wT_,
                                const void *bT_, const void *eaP_, const void *eaN_, const void *aD_, void *out,
                                void *outm)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const T *chT = (const T *)chT_, *shT = (const T *)shT_, *wT = (const T *)wT_, *bT = (const T *)bT_;
    const PadLayout L = pad_layout(CPLX, N, K, Mp);
    const long long KN = (long long)K * N, CSR = Mp + 1;
    D *d = (D *)out;
    ChSh<T> *cs = (ChSh<T> *)((unsigned char *)out + L.cs);
    T *w = (T *)((unsigned char *)out + L.w);
    T *b = (T *)((unsigned char *)out + L.b);
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    ChSh<T> *csm = (ChSh<T> *)((unsigned char *)outm + L.cs);
    T *wm = (T *)((unsigned char *)outm + L.w);
    T *bm = (T *)((unsigned char *)outm + L.b);
    D *dm = (D *)outm;
    for (long long q = t0; q < KN * CSR; q += stride) {
        const long long row = q / CSR;
        const int j = (int)(q % CSR);
        ChSh<T> v;
        v.ch = j < M ? chT[row * M + j] : oneT<T>();
        v.sh = j < M ? shT[row * M + j] : zeroT<T>();
        cs[q] = v;
        if (j < Mp) csm[((row / N) * Mp + j) * N + row % N] = v;  // row = kk N + i
    }
    for (long long q = t0; q < KN * Mp; q += stride) {
        const long long row = q / Mp;
        const int j = (int)(q % Mp);
        w[q] = wm[q] = j < M ? wT[row * M + j] : zeroT<T>();
    }
    for (long long q = t0; q < (long long)K * Mp; q += stride) {
        const int kk = (int)(q / Mp), j = (int)(q % Mp);
        b[q] = bm[q] = j < M ? bT[(long long)kk * M + j] : zeroT<T>();
    }
    for (long long q = t0; q < KN; q += stride) {
        d[q] = dm[q] = ((const D *)eaP_)[q];
        d[KN + q] = dm[KN + q] = ((const D *)eaN_)[q];
        d[2 * KN + q] = dm[2 * KN + q] = ((const D *)aD_)[q];
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
                    // theta +- w as one FMA with the spin's sign (exact product by +-1)
                    const float sg = ((bits[k] >> ii) & 1u) ? -1.0f : 1.0f;
#pragma unroll
                    for (int kk = 0; kk < K; kk++)
#pragma unroll
                        for (int q = 0; q < JQ; q++) th[k][kk][q] = addsT(th[k][kk][q], sg, wv[kk][q]);
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

// Metropolis chains (see the file comment).  Tables in LDS when TAB, else the cosh/sinh
// rows read from the padded global copy (L2-resident) and exp(+-2a), a staged in LDS.  Replica k's angles and Phi row are held in
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
    const PadLayout PL = pad_layout(CPLX, N, K, Mp);
    const D *eaP = dtab, *eaN = dtab + KN, *aD = dtab + 2 * KN;
    if constexpr (!TAB) {  // exp(+-2a) and a in LDS: read every flip
        D *le = (D *)smem;
        for (long long q = tid; q < 3 * KN; q += blockDim.x) le[q] = dtab[q];
        eaP = le;
        eaN = le + KN;
        aD = le + 2 * KN;
        p = smem + ((3 * KN * sizeof(D) + 15) & ~(size_t)15);
    }
    const ChSh<T> *cs = (const ChSh<T> *)((const unsigned char *)dtab + PL.cs);
    const T *wp = (const T *)((const unsigned char *)dtab + PL.w);
    const T *bp = (const T *)((const unsigned char *)dtab + PL.b);
    constexpr int CSR = Mp + 1;
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
            uint32_t cur = sp[k * NW];
            for (int i = 0; i < N; i++) {
                ChSh<T> cv[K][JQ];
                if constexpr (TAB) {
                    // row (kk, i) = row (0, i) + kk N CSR: one lane address per flip, the
                    // rest constant offsets
                    const ChSh<T> *r0 = cs + (i * CSR + g);
#pragma unroll
                    for (int kk = 0; kk < K; kk++)
#pragma unroll
                        for (int q = 0; q < JQ; q++) cv[kk][q] = r0[kk * N * CSR + G * q];
                } else {
                    // from L2 at the top of the flip (a one-flip-ahead prefetch into registers
                    // held 2 K JQ more and, at K = 4, halved the wavefronts per SIMD: slower)
#pragma unroll
                    for (int kk = 0; kk < K; kk++)
#pragma unroll
                        for (int q = 0; q < JQ; q++) cv[kk][q] = cs[((long long)kk * N + i) * CSR + g + G * q];
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
                    // K <= 2: VALU issue bounds the sweeps (2 wavefronts per SIMD), so the
                    // cross-row exchanges go through the LDS crossbar; with more
                    // replicas the register-heavy kernel runs fewer wavefronts and the
                    // LDS latency would show
                    const D Tp = group_prod_tf<G, T, D, (K <= 2)>(pf[kk]);
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
                // draw tdraw of each group's batch: lane grp G + tdraw's value, fetched
                // through the LDS crossbar (no VALU work but the address)
                double u;
                {
                    const int addr = (grp * G + tdraw) << 2;
                    const uint64_t ubits = __builtin_bit_cast(uint64_t, ub);
                    const uint32_t lo32 = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)ubits);
                    const uint32_t hi32 = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(ubits >> 32));
                    u = __builtin_bit_cast(double, (uint64_t)lo32 | ((uint64_t)hi32 << 32));
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

// ---------------------------------------------------------------- measurement
// hs-like local energy (rbm.cc:1092-1119) of one sample held by a group of G lanes:
// res row k = sum over pairs i > j of J_ij (2 Phi^(ij)_k - delta Phi_k) if
// s^k_i != s^k_j, else J_ij delta Phi_k; H = Phi^-1 res.  Phi^(ij)_k flips spin i,
// then spin j of replica k (the second flip on the angles after the first,
// rbm.cc:1106-1108).  Lane g takes the pairs with index = g mod G; pairs with
// J_ij == 0 are skipped.  Products over hidden units in T, flushed to D every 8.
template <bool CPLX, int K, int G, int Mp>
__device__ inline void hloc_pairflip(const MixedSampleArgs &a, const ChSh<typename Ty<CPLX>::T> *cs,
                                     const typename Ty<CPLX>::D *eaP, const typename Ty<CPLX>::D *eaN,
                                     const double *Jm, const uint32_t *sp, const typename Ty<CPLX>::T *tl, int g,
                                     const typename Ty<CPLX>::D *phi, const typename Ty<CPLX>::D *phi_inv,
                                     typename Ty<CPLX>::D *H)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const int N = a.N, M = a.M, NW = a.NW;
    const double delta = a.delta;
    D res[K * K];
#pragma unroll
    for (int q = 0; q < K * K; q++) res[q] = zeroD<D>();
    int pr = 0;
    for (int i = 1; i < N; i++)
        for (int j = 0; j < i; j++, pr++) {
            if (pr % G != g) continue;
            const double pij = Jm[(long long)i * N + j];
            if (pij == 0.0) continue;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int si = (sp[k * NW + (i >> 5)] >> (i & 31)) & 1u;
                const int sj = (sp[k * NW + (j >> 5)] >> (j & 31)) & 1u;
                if (si == sj) {
#pragma unroll
                    for (int kk = 0; kk < K; kk++) res[k * K + kk] = addD(res[k * K + kk], scaleD(pij * delta, phi[k * K + kk]));
                    continue;
                }
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    const ChSh<T> *ri = cs + (long long)kk * Mp * N + i, *rj = cs + (long long)kk * Mp * N + j;
                    const T *tk = tl + (k * K + kk) * Mp;
                    D prod = oneD<D>();
                    T pf = oneT<T>();
                    for (int hj = 0; hj < M; hj++) {
                        const T t = tk[hj];
                        const ChSh<T> ci = ri[(long long)hj * N], cj = rj[(long long)hj * N];
                        const T s_i = si ? negT(ci.sh) : ci.sh;
                        const T den1 = subT(ci.ch, mulT(t, s_i));
                        const T t1 = divT(subT(mulT(t, ci.ch), s_i), den1);
                        const T s_j = sj ? negT(cj.sh) : cj.sh;
                        const T den2 = subT(cj.ch, mulT(t1, s_j));
                        pf = mulT(pf, mulT(den1, den2));
                        if ((hj & 7) == 7) {
                            prod = mulD(prod, toD(pf));
                            pf = oneT<T>();
                        }
                    }
                    prod = mulD(prod, toD(pf));
                    const D e = mulD(si ? eaP[kk * N + i] : eaN[kk * N + i], sj ? eaP[kk * N + j] : eaN[kk * N + j]);
                    const D v = mulD(mulD(phi[k * K + kk], e), prod);
                    res[k * K + kk] = addD(res[k * K + kk], scaleD(pij, subD(scaleD(2.0, v), scaleD(delta, phi[k * K + kk]))));
                }
            }
        }
#pragma unroll
    for (int q = 0; q < K * K; q++) res[q] = group_sum<G>(res[q]);
    matmul_m<K, D>(phi_inv, res, H);
}

// Local energy and factor rows of sample s from its state (sp, tl [k][kk][Mp], phi)
// held by a group of G lanes: E_zz (rbm.cc:1012-1031), H_loc = Phi^-1 (diag(E_zz) Phi
// - h sum_i Phi^(i)) (rbm.cc:1042-1064) or the pair-flip H_loc above, and
// O_{k,p} = sum_kk c_{k,kk} d_{kk,p} with c_{k,kk} = (Phi^-1)_{k,kk} Phi_{kk,k}
// (Cramer's rule applied to rbm.cc:199-239) as the O row or the factor rows U.
// sum_i Phi^(i): lanes take spins i = g, g + G, ... two at a time and run over all
// hidden units j; every cosh/sinh pair read feeds the K replicas and every angle read
// (the same address for the whole group: an LDS broadcast) feeds both spins.
template <bool CPLX, int K, int G, int Mp>
__device__ inline void measure_sample(const MixedSampleArgs &a, const ChSh<typename Ty<CPLX>::T> *cs,
                                      const typename Ty<CPLX>::D *eaP, const typename Ty<CPLX>::D *eaN,
                                      const double *Jm, const uint32_t *sp, const typename Ty<CPLX>::T *tl, int g,
                                      const typename Ty<CPLX>::D *phi, long long s, bool live)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const int N = a.N, M = a.M, NW = a.NW, L = a.L;
    D phi_inv[K * K];
    inv_m<K, D>(phi, phi_inv);

    D H[K * K];
    if (a.model != 0) {
        hloc_pairflip<CPLX, K, G, Mp>(a, cs, eaP, eaN, Jm, sp, tl, g, phi, phi_inv, H);
    } else {
        // E_zz per replica: lanes take rows i = g, g+G, ..., and the columns j < i of
        // row i starting at a lane-dependent offset (rows of J share LDS banks)
        double ez[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            double r = 0.0;
            for (int i = g; i < N; i += G) {
                const int bi = (sp[k * NW + (i >> 5)] >> (i & 31)) & 1u;
                const double *Ji = Jm + (long long)i * N;
                int j = g % N;
                for (int jj = 0; jj < N; jj++) {
                    if (j < i) {
                        const int bj = (sp[k * NW + (j >> 5)] >> (j & 31)) & 1u;
                        const double v = Ji[j];
                        r += (bi ^ bj) ? -v : v;
                    }
                    if (++j == N) j = 0;
                }
            }
            ez[k] = 2.0 * group_sum<G>(r);
        }
        D sx[K * K];
#pragma unroll
        for (int q = 0; q < K * K; q++) sx[q] = zeroD<D>();
        // the terms of spins i0 and (TWO) i1 = i0 + G of replica kk, added to sx in that order
        auto spin_terms = [&](int kk, int i0, auto two_c) {
            constexpr bool TWO = decltype(two_c)::value;
            const int i1 = TWO ? i0 + G : i0;
            bool dn0[K], dn1[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                dn0[k] = (sp[k * NW + (i0 >> 5)] >> (i0 & 31)) & 1u;
                dn1[k] = (sp[k * NW + (i1 >> 5)] >> (i1 & 31)) & 1u;
            }
            const ChSh<T> *r0 = cs + (long long)kk * Mp * N + i0, *r1 = cs + (long long)kk * Mp * N + i1;
            D prod0[K], prod1[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                prod0[k] = oneD<D>();
                prod1[k] = oneD<D>();
            }
#pragma unroll 1
            for (int j0 = 0; j0 < Mp; j0 += 8) {
                T pf0[K], pf1[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    pf0[k] = oneT<T>();
                    pf1[k] = oneT<T>();
                }
#pragma unroll
                for (int jj = 0; jj < 8; jj++) {
                    const ChSh<T> c0 = r0[(long long)(j0 + jj) * N];
                    ChSh<T> c1;
                    if constexpr (TWO) c1 = r1[(long long)(j0 + jj) * N];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const T t = tl[(k * K + kk) * Mp + j0 + jj];
                        pf0[k] = mulT(pf0[k], subT(c0.ch, mulT(t, dn0[k] ? negT(c0.sh) : c0.sh)));
                        if constexpr (TWO) pf1[k] = mulT(pf1[k], subT(c1.ch, mulT(t, dn1[k] ? negT(c1.sh) : c1.sh)));
                    }
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    prod0[k] = mulD(prod0[k], toD(pf0[k]));
                    if constexpr (TWO) prod1[k] = mulD(prod1[k], toD(pf1[k]));
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const D e0 = dn0[k] ? eaP[kk * N + i0] : eaN[kk * N + i0];
                sx[k * K + kk] = addD(sx[k * K + kk], mulD(mulD(phi[k * K + kk], e0), prod0[k]));
                if constexpr (TWO) {
                    const D e1 = dn1[k] ? eaP[kk * N + i1] : eaN[kk * N + i1];
                    sx[k * K + kk] = addD(sx[k * K + kk], mulD(mulD(phi[k * K + kk], e1), prod1[k]));
                }
            }
        };
        // lane g takes spins g, g + G, g + 2G, ...: two at a time while two remain (every
        // angle read feeds both), then a last one alone (N <= G: one spin per lane)
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            int i0 = g;
            for (; i0 + G < N; i0 += 2 * G) spin_terms(kk, i0, std::true_type{});
            if (i0 < N) spin_terms(kk, i0, std::false_type{});
        }
        D res[K * K];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                const D sxs = group_sum<G>(sx[k * K + kk]);
                res[k * K + kk] = subD(scaleD(ez[k], phi[k * K + kk]), scaleD(a.h, sxs));
            }
        matmul_m<K, D>(phi_inv, res, H);
    }
    D energy = zeroD<D>();
#pragma unroll
    for (int k = 0; k < K; k++) energy = addD(energy, H[k * K + k]);
    if (live) {
        cd *Hs = (cd *)a.Hs;
        cd *Es = (cd *)a.Es;
#pragma unroll
        for (int q = 0; q < K * K; q++)
            if (g == q) {
                cd hv;
                if constexpr (CPLX) hv = H[q];
                else hv = nq64::mk(H[q], 0.0);
                Hs[s * K * K + q] = hv;
            }
        if (g == 0) {
            cd ev;
            if constexpr (CPLX) ev = energy;
            else ev = nq64::mk(energy, 0.0);
            Es[s] = ev;
        }
    }

    // O row: O_{k,p} = sum_kk c_{k,kk} d_{kk,p}; factor rows U[s K + kk]
    if (live && (a.O || a.U)) {
        T cT[K * K];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int kk = 0; kk < K; kk++) cT[k * K + kk] = toT(mulD(phi_inv[k * K + kk], phi[kk * K + k]));
        T *Orow = a.O ? (T *)a.O + s * a.Pld : nullptr;
        T *Ub = a.U ? (T *)a.U + s * K * (long long)a.Ucols : nullptr;
        float umx[K];  // real weights: this lane's max |U| of each of the sample's K rows
#pragma unroll
        for (int kk = 0; kk < K; kk++) umx[kk] = 0.f;
#pragma unroll
        for (int k = 0; k < K; k++) {
            T ck[K];
#pragma unroll
            for (int kk = 0; kk < K; kk++) ck[kk] = cT[k * K + kk];
            T *ob = Orow + (long long)k * L;
            // a block: sum_kk c sigma_i^kk
            if (Orow)
                for (int i = g; i < N; i += G) {
                    T v = zeroT<T>();
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        int bit = (sp[kk * NW + (i >> 5)] >> (i & 31)) & 1u;
                        v = bit ? subT(v, ck[kk]) : addT(v, ck[kk]);
                    }
                    ob[i] = v;
                }
            // b block and w block: lanes over j
            for (int j = g; j < M; j += G) {
                T ct[K];
                T v = zeroT<T>();
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    ct[kk] = mulT(ck[kk], tl[(kk * K + k) * Mp + j]);
                    v = addT(v, ct[kk]);
                }
                if (Ub)
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        Ub[(long long)kk * a.Ucols + k * M + j] = ct[kk];
                        if constexpr (!CPLX) umx[kk] = fmaxf(umx[kk], fabsf(ct[kk]));
                    }
                if (!Orow) continue;
                ob[N + j] = v;
                for (int i = 0; i < N; i++) {
                    T w = zeroT<T>();
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        int bit = (sp[kk * NW + (i >> 5)] >> (i & 31)) & 1u;
                        w = bit ? subT(w, ct[kk]) : addT(w, ct[kk]);
                    }
                    ob[N + M + (long long)i * M + j] = w;
                }
            }
        }
        if (Ub)
            for (int col = K * M + g; col < a.Ucols; col += G)
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    T val = zeroT<T>();
#pragma unroll
                    for (int k = 0; k < K; k++)
                        if (col - K * M == k) val = cT[k * K + kk];
                    Ub[(long long)kk * a.Ucols + col] = val;
                }
        if constexpr (!CPLX) {
            // max |U| of each of the sample's K rows (the int8 form's row scales): the
            // lane's entries and the c columns, then the G lanes of the sample (exact: a
            // maximum)
            if (a.Umax && Ub) {
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    float mx = umx[kk];
#pragma unroll
                    for (int k = 0; k < K; k++) mx = fmaxf(mx, fabsf(cT[k * K + kk]));
#pragma unroll
                    for (int off = 1; off < G; off <<= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
                    if (g == 0) a.Umax[s * K + kk] = mx;
                }
            }
        }
        if (Orow)
            for (long long q = a.P + g; q < a.Pld; q += G) Orow[q] = zeroT<T>();
    }
}

// Wavefronts per SIMD the measurement kernel is compiled for: 2 (at most 256 VGPRs) for
// the real-weight shapes whose unconstrained build lands just above 256 registers (K = 4
// with up to 4 hidden units per lane, K = 3 with 64 lanes and 8 units per lane), which
// otherwise run one wavefront per SIMD; capped they spill at most 2 registers.  1 (no
// bound) elsewhere.
__host__ __device__ constexpr int measure_min_waves(bool cplx, int K, int G, int JQ)
{
    return !cplx && ((K == 4 && JQ <= 4) || (K == 3 && G == 64 && JQ == 8)) ? 2 : 1;
}

// Local energies and O rows (or factor rows) of all S samples from their stored
// spins: G lanes per sample, 64/G samples per wavefront, each block looping over
// sample batches (grid sized to the CUs) so that the tables are staged in LDS once
// per block.  Angles and Phi from the spins as in the sweeps (sweep_refresh, lanes
// over hidden units), then measure_sample.
template <bool CPLX, int K, int G, int JQ, bool TAB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(measure_min_waves(CPLX, K, G, JQ))))
void k_measure_mx(MixedSampleArgs a, const void *tabp)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    constexpr int SPW = 64 / G;
    constexpr int Mp = G * JQ;
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, NW = a.NW;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane & (G - 1), grp = lane / G;
    const int spb = (blockDim.x >> 6) * SPW;
    const int sib = wv * SPW + grp;
    const long long S = (long long)a.C * a.m;
    const long long KN = (long long)K * N;

    const D *dtab;
    const double *Jm = a.J;
    unsigned char *p = smem;
    const PadLayout PL = pad_layout(CPLX, N, K, Mp);
    if constexpr (TAB) {
        const uint4 *src = (const uint4 *)tabp;
        uint4 *dst = (uint4 *)smem;
        for (size_t q = tid; q < PL.total / 16; q += blockDim.x) dst[q] = src[q];
        double *jl = (double *)(smem + PL.total);
        for (int q = tid; q < N * N; q += blockDim.x) jl[q] = a.J[q];
        dtab = (const D *)smem;
        Jm = jl;
        p = smem + al16(PL.total + (size_t)N * N * sizeof(double));
    } else {
        dtab = (const D *)tabp;
    }
    const D *eaP = dtab, *eaN = dtab + KN, *aD = dtab + 2 * KN;
    if constexpr (!TAB) {  // exp(+-2a), a and (when a.jlds) J in LDS; the cosh/sinh, w, b tables from L2
        D *le = (D *)smem;
        for (long long q = tid; q < 3 * KN; q += blockDim.x) le[q] = dtab[q];
        eaP = le;
        eaN = le + KN;
        aD = le + 2 * KN;
        p = smem + al16((size_t)3 * KN * sizeof(D));
        if (a.jlds) {
            double *jl = (double *)p;
            for (int q = tid; q < N * N; q += blockDim.x) jl[q] = a.J[q];
            Jm = jl;
            p += al16((size_t)N * N * sizeof(double));
        }
    }
    const ChSh<T> *cs = (const ChSh<T> *)((const unsigned char *)dtab + PL.cs);
    const T *wp = (const T *)((const unsigned char *)dtab + PL.w);
    const T *bp = (const T *)((const unsigned char *)dtab + PL.b);
    unsigned char *st = p + (size_t)sib * chain_lds_bytes<CPLX, K, G, JQ>(NW);
    D *phis = (D *)st;
    T *tls = (T *)(phis + K * K);
    uint32_t *sp = (uint32_t *)(tls + K * K * Mp);
    __syncthreads();
    int overflow = 0;
    for (long long base = (long long)blockIdx.x * spb; base < S; base += (long long)gridDim.x * spb) {
        const long long s_raw = base + sib;
        const bool live = s_raw < S;
        const long long s = live ? s_raw : S - 1;
        if (g == 0)
            for (int w = 0; w < K * NW; w++) sp[w] = a.spins[s * K * NW + w];
        __builtin_amdgcn_wave_barrier();
        sweep_refresh<CPLX, K, G, JQ>(a, aD, wp, bp, sp, tls, phis, g, 0, false, overflow);
        D phi[K * K];
#pragma unroll
        for (int q = 0; q < K * K; q++) phi[q] = phis[q];
        measure_sample<CPLX, K, G, Mp>(a, cs, eaP, eaN, Jm, sp, tls, g, phi, s, live);
        __builtin_amdgcn_wave_barrier();
    }
    if (overflow) atomicOr(a.flags, 1u);
}

template <bool CPLX, int K, int G, int JQ>
hipError_t launch_measure_jq(const MixedSampleArgs &a, const void *tabp, const MeasureShape &sh, hipStream_t st)
{
    const int spb = sh.wpb * (64 / G);
    size_t lds = (size_t)spb * chain_lds_bytes<CPLX, K, G, JQ>(a.NW);
    if (sh.tab) lds += al16(pad_tab_bytes(CPLX, a.N, K, G * JQ) + (size_t)a.N * a.N * sizeof(double));
    else lds += al16((size_t)3 * K * a.N * (CPLX ? 16 : 8)) + (a.jlds ? al16((size_t)a.N * a.N * sizeof(double)) : 0);
    const long long S = (long long)a.C * a.m;
    long long blocks = (S + spb - 1) / spb;
    const void *fn = sh.tab ? (const void *)k_measure_mx<CPLX, K, G, JQ, true> : (const void *)k_measure_mx<CPLX, K, G, JQ, false>;
    if (lds > 65536) {
        hipError_t e = nq_allow_dynamic_lds(fn, lds);
        if (e != hipSuccess) return e;
    }
    // resident blocks: limited by LDS (160 KB per CU) and by 8 waves per SIMD
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    long long per_cu = lds > 0 ? (long long)(160 * 1024) / (long long)lds : 8;
    if (per_cu > 32 / sh.wpb) per_cu = 32 / sh.wpb;
    if (per_cu < 1) per_cu = 1;
    const char *eb = getenv("NQES_MEASURE_BLOCKS_PER_CU");
    if (eb && *eb && atoi(eb) > 0) per_cu = atoi(eb);
    if (blocks > per_cu * ncu) blocks = per_cu * ncu;
    if (sh.tab) hipLaunchKernelGGL((k_measure_mx<CPLX, K, G, JQ, true>), dim3((unsigned)blocks), dim3(64 * sh.wpb), lds, st, a, tabp);
    else hipLaunchKernelGGL((k_measure_mx<CPLX, K, G, JQ, false>), dim3((unsigned)blocks), dim3(64 * sh.wpb), lds, st, a, tabp);
    return hipGetLastError();
}

template <bool CPLX, int K>
hipError_t launch_measure_k(const MixedSampleArgs &a, const void *tabp, const MeasureShape &sh, hipStream_t st)
{
    // the shapes nq_measure_shape picks for this K
    const int key = sh.G * 16 + sh.JQ;
    if constexpr (K <= 2) {
        switch (key) {
        case 16 * 16 + 2: return launch_measure_jq<CPLX, K, 16, 2>(a, tabp, sh, st);
        case 16 * 16 + 4: return launch_measure_jq<CPLX, K, 16, 4>(a, tabp, sh, st);
        case 16 * 16 + 8: return launch_measure_jq<CPLX, K, 16, 8>(a, tabp, sh, st);
        case 32 * 16 + 8: return launch_measure_jq<CPLX, K, 32, 8>(a, tabp, sh, st);
        case 64 * 16 + 8: return launch_measure_jq<CPLX, K, 64, 8>(a, tabp, sh, st);
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (key) {
        case 16 * 16 + 2: return launch_measure_jq<CPLX, K, 16, 2>(a, tabp, sh, st);
        case 32 * 16 + 2: return launch_measure_jq<CPLX, K, 32, 2>(a, tabp, sh, st);
        case 64 * 16 + 2: return launch_measure_jq<CPLX, K, 64, 2>(a, tabp, sh, st);
        case 64 * 16 + 4: return launch_measure_jq<CPLX, K, 64, 4>(a, tabp, sh, st);
        case 64 * 16 + 8: return launch_measure_jq<CPLX, K, 64, 8>(a, tabp, sh, st);
        default: return hipErrorInvalidValue;
        }
    }
}

template <bool CPLX, int K, int G, int JQ>
hipError_t launch_jq(const MixedSampleArgs &a, const void *tabp, bool tab, int wpb, hipStream_t st)
{
    const int cpb = wpb * (64 / G);
    size_t lds = (size_t)cpb * chain_lds_bytes<CPLX, K, G, JQ>(a.NW);
    if (tab) lds += pad_tab_bytes(CPLX, a.N, K, G * JQ);
    else lds += ((size_t)3 * K * a.N * (CPLX ? 16 : 8) + 15) & ~(size_t)15;
    const int blocks = (a.C + cpb - 1) / cpb;
    const void *fn = tab ? (const void *)k_sample_mx<CPLX, K, G, JQ, true> : (const void *)k_sample_mx<CPLX, K, G, JQ, false>;
    if (lds > 65536) {
        hipError_t e = nq_allow_dynamic_lds(fn, lds);
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

template <>
hipError_t nq_measure_k<NQES_SWEEP_K>(bool cplx, const MixedSampleArgs &a, const void *tabp, const MeasureShape &sh,
                                      hipStream_t st)
{
    return cplx ? launch_measure_k<true, NQES_SWEEP_K>(a, tabp, sh, st)
                : launch_measure_k<false, NQES_SWEEP_K>(a, tabp, sh, st);
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
    const size_t ea = al16((size_t)3 * K * N * (cplx ? 16 : 8));  // staged when the tables stay in L2
    if (lds_max > 0) {
        while (wpb > 1 && states(wpb) + ea > lds_max) wpb--;
        if (states(wpb) + ea > lds_max) return sh;
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

// Measurement: the fewest lanes per sample (16, 32, 64) that keep the refresh's
// K*K*JQ angles per lane (JQ = Mp / G) at <= 32 with JQ <= 8, else 64 lanes; tables
// and J in LDS when they fit next to the per-sample states of 4 wavefronts.
MeasureShape nq_measure_shape(bool cplx, int N, int K, const SweepShape &sw, size_t lds_max)
{
    MeasureShape sh{};
    if (!sw.ok) return sh;
    const int Mp = sw.G * sw.JQ;
    sh.G = 64;
    for (int G = 16; G <= 64; G *= 2)
        if (Mp / G <= 8 && K * K * (Mp / G) <= 32) {
            sh.G = G;
            break;
        }
    sh.JQ = Mp / sh.G;
    sh.wpb = 4;
    const int NW = (N + 31) / 32;
    const size_t per_sample =
        (size_t)K * K * Mp * (cplx ? 8 : 4) + (size_t)K * K * (cplx ? 16 : 8) + (size_t)K * NW * 4;
    const size_t states = (size_t)sh.wpb * (64 / sh.G) * per_sample;
    if (lds_max > 0 && states > lds_max) return sh;
    const size_t tb = al16(pad_tab_bytes(cplx, N, K, Mp) + (size_t)N * N * sizeof(double));
    const char *et = getenv("NQES_TABLDS");
    sh.tab = (tb + states <= lds_max) && !(et && *et && atoi(et) == 0);
    // tables in L2: exp(+-2a) and a in LDS, and J when it fits too
    const size_t ea = al16((size_t)3 * K * N * (cplx ? 16 : 8));
    if (!sh.tab && lds_max > 0 && states + ea > lds_max) return sh;
    sh.jlds = !sh.tab && (lds_max == 0 || states + ea + al16((size_t)N * N * sizeof(double)) <= lds_max);
    sh.ok = 1;
    return sh;
}

hipError_t nq_launch_measure_mx(bool cplx, int K, const MixedSampleArgs &a, const MeasureShape &sh, const void *tabp,
                                hipStream_t st)
{
    if (!sh.ok) return hipErrorInvalidValue;
    MixedSampleArgs b = a;
    b.jlds = sh.jlds;
    switch (K) {
    case 1: return nq_measure_k<1>(cplx, b, tabp, sh, st);
    case 2: return nq_measure_k<2>(cplx, b, tabp, sh, st);
    case 3: return nq_measure_k<3>(cplx, b, tabp, sh, st);
    case 4: return nq_measure_k<4>(cplx, b, tabp, sh, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t nq_launch_sweep_mx(bool cplx, int K, const MixedSampleArgs &a, const SweepShape &sh, void *tabp,
                              void *tabm, hipStream_t st)
{
    if (!sh.ok) return hipErrorInvalidValue;
    const int Mp = sh.G * sh.JQ;
    const long long n = (long long)K * a.N * Mp;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (cplx)
        hipLaunchKernelGGL(k_pad_tables_mx<true>, dim3(blocks), dim3(256), 0, st, a.N, a.M, K, Mp, a.chT, a.shT, a.wT,
                           a.bT, a.eaP, a.eaN, a.aD, tabp, tabm);
    else
        hipLaunchKernelGGL(k_pad_tables_mx<false>, dim3(blocks), dim3(256), 0, st, a.N, a.M, K, Mp, a.chT, a.shT,
                           a.wT, a.bT, a.eaP, a.eaN, a.aD, tabp, tabm);
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
