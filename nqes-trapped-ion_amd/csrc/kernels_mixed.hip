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
//   (k_sample_mx, the Metropolis sweeps, is in sampler_mx.hip)
//   k_measure_mx     G lanes per sample: theta, tanh, Phi, Phi^-1 from the stored spins,
//                    local energy (rbm.cc:1042-1064) and the O row
//                    O_{k,p} = sum_kk c_{k,kk} d_{kk,p} with c_{k,kk} = (Phi^-1)_{k,kk} Phi_{kk,k}
//                    (Cramer's rule applied to rbm.cc:199-239) written to HBM in FP32
//   k_moments_mx     column sums of O, |O|^2, E conj(O) in FP64 with Neumaier compensation
//   k_sv_rows_mx     S*v in one streamed pass over O (rbm.cc:661-684): WPR wavefronts
//                    share a row, w_s = (O_s - o)^T v, then acc += conj(O_s) w_s with
//                    FP64 accumulation in registers; next row prefetched, nontemporal loads
//   k_sv_rowdot_mx / k_sv_colacc_mx   two-pass S*v when v does not fit in LDS
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "cplx64.hpp"
#include "kernels.hpp"
#include "mixed_math.hpp"

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

// ---------------------------------------------------------------- sampler
// Per-step tables read by the sampler and the measurement: cosh(2w), sinh(2w), w
// ([kk][i][j]), b ([kk][j]) in T and exp(2a), exp(-2a), a ([kk][i]) in D; in LDS when
// the kernel's TAB parameter (mx_tab_bytes), else in global memory.
template <bool CPLX> struct MxTab {
    const typename Ty<CPLX>::T *ch, *sh, *w, *b;
    const typename Ty<CPLX>::D *eaP, *eaN, *aD;
};

__host__ __device__ inline size_t mx_tab_bytes(bool cplx, int N, int M, int K)
{
    const size_t ts = cplx ? 8 : 4, ds = cplx ? 16 : 8;
    return 3 * (size_t)K * N * ds + (3 * (size_t)K * N * M + (size_t)K * M) * ts;
}

// point tb at the tables (copied into LDS at p when TAB); returns the LDS
// cursor after them.  D arrays first to keep their alignment.
template <bool CPLX, bool TAB>
__device__ inline unsigned char *mx_stage_tables(const MixedSampleArgs &a, int K, MxTab<CPLX> &tb, unsigned char *p)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    tb.ch = (const T *)a.chT;
    tb.sh = (const T *)a.shT;
    tb.w = (const T *)a.wT;
    tb.b = (const T *)a.bT;
    tb.eaP = (const D *)a.eaP;
    tb.eaN = (const D *)a.eaN;
    tb.aD = (const D *)a.aD;
    if constexpr (!TAB) return p;
    const long long KN = (long long)K * a.N, KNM = KN * a.M, KM = (long long)K * a.M;
    D *d = (D *)p;
    for (long long q = threadIdx.x; q < KN; q += blockDim.x) {
        d[q] = tb.eaP[q];
        d[KN + q] = tb.eaN[q];
        d[2 * KN + q] = tb.aD[q];
    }
    T *t = (T *)(d + 3 * KN);
    for (long long q = threadIdx.x; q < KNM; q += blockDim.x) {
        t[q] = tb.ch[q];
        t[KNM + q] = tb.sh[q];
        t[2 * KNM + q] = tb.w[q];
    }
    for (long long q = threadIdx.x; q < KM; q += blockDim.x) t[3 * KNM + q] = tb.b[q];
    tb.eaP = d;
    tb.eaN = d + KN;
    tb.aD = d + 2 * KN;
    tb.ch = t;
    tb.sh = t + KNM;
    tb.w = t + 2 * KNM;
    tb.b = t + 3 * KNM;
    return (unsigned char *)(t + 3 * KNM + KM);
}
template <typename T>
__device__ inline T ld(const T *p)
{
    return *p;
}

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

// Theta -> tanh (tl, LDS) and Phi for one chain state held by a group of G lanes.
// Lane g owns hidden units j = g + G q.  Source: Theta0 of chain c (chain start) or
// the spins sp (LDS).  Phi_{k,kk} = exp(sum_i +-a_i) prod_j cosh(theta_j) with the
// cosh product taken per lane in FP64, then across the group.
template <bool CPLX, int K, int G, int JMAX>
__device__ inline void mx_refresh(const MixedSampleArgs &a, const MxTab<CPLX> &tb, const uint32_t *sp,
                                  typename Ty<CPLX>::T *tl, int g, long long c, bool from_theta0,
                                  typename Ty<CPLX>::D *phi, int &overflow)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const int N = a.N, M = a.M, NW = a.NW;
    const int jpl = (M + G - 1) / G;
    const T *wT = tb.w;
    const T *bT = tb.b;
    const D *aD = tb.aD;
    D part[K * K];
#pragma unroll
    for (int q = 0; q < K * K; q++) part[q] = oneD<D>();
#pragma unroll
    for (int q = 0; q < JMAX; q++) {
        int j = g + G * q;
        if (q < jpl && j < M) {
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                T th[K];
                if (from_theta0) {
#pragma unroll
                    for (int k = 0; k < K; k++) th[k] = ((const T *)a.theta0)[((c * K + k) * K + kk) * M + j];
                } else {
                    T b = bT[kk * M + j];
#pragma unroll
                    for (int k = 0; k < K; k++) th[k] = b;
                    for (int i = 0; i < N; i++) {
                        T w = wT[((long long)kk * N + i) * M + j];
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            int bit = (sp[k * NW + (i >> 5)] >> (i & 31)) & 1u;
                            th[k] = bit ? subT(th[k], w) : addT(th[k], w);
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    if (fabsf(reT(th[k])) > 80.0f) overflow = 1;
                    tl[(k * K + kk) * M + j] = tanhT(th[k]);
                    part[k * K + kk] = mulD(part[k * K + kk], toD(coshT(th[k])));
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            D A = zeroD<D>();
            for (int i = 0; i < N; i++) {
                int bit = (sp[k * NW + (i >> 5)] >> (i & 31)) & 1u;
                D ai = aD[kk * N + i];
                A = bit ? subD(A, ai) : addD(A, ai);
            }
            phi[k * K + kk] = mulD(expD(A), group_prod<G>(part[k * K + kk]));
        }
}

// hs-like local energy (rbm.cc:1092-1119) of one sample held by a group of G lanes:
// res row k = sum over pairs i > j of J_ij (2 Phi^(ij)_k - delta Phi_k) if
// s^k_i != s^k_j, else J_ij delta Phi_k; H = Phi^-1 res.  Phi^(ij)_k flips spin i,
// then spin j of replica k (the second flip on the angles after the first,
// rbm.cc:1106-1108).  Lane g takes the pairs with index = g mod G; pairs with
// J_ij == 0 are skipped.  Products over hidden units in T, flushed to D every 8.
template <bool CPLX, int K, int G>
__device__ inline void mx_hloc_hs(const MixedSampleArgs &a, const MxTab<CPLX> &tb, const double *Jm,
                                  const uint32_t *sp, const typename Ty<CPLX>::T *tl, int g,
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
                    const T *chi = tb.ch + ((long long)kk * N + i) * M, *shi = tb.sh + ((long long)kk * N + i) * M;
                    const T *chj = tb.ch + ((long long)kk * N + j) * M, *shj = tb.sh + ((long long)kk * N + j) * M;
                    const T *tk = tl + (k * K + kk) * M;
                    D prod = oneD<D>();
                    T pf = oneT<T>();
                    for (int hj = 0; hj < M; hj++) {
                        const T t = tk[hj];
                        const T ci = chi[hj], s_i = si ? negT(shi[hj]) : shi[hj];
                        const T den1 = subT(ci, mulT(t, s_i));
                        const T t1 = divT(subT(mulT(t, ci), s_i), den1);
                        const T s_j = sj ? negT(shj[hj]) : shj[hj];
                        const T den2 = subT(chj[hj], mulT(t1, s_j));
                        pf = mulT(pf, mulT(den1, den2));
                        if ((hj & 7) == 7) {
                            prod = mulD(prod, toD(pf));
                            pf = oneT<T>();
                        }
                    }
                    prod = mulD(prod, toD(pf));
                    const D e = mulD(si ? tb.eaP[kk * N + i] : tb.eaN[kk * N + i], sj ? tb.eaP[kk * N + j] : tb.eaN[kk * N + j]);
                    const D v = mulD(mulD(phi[k * K + kk], e), prod);
                    res[k * K + kk] = addD(res[k * K + kk], scaleD(pij, subD(scaleD(2.0, v), scaleD(delta, phi[k * K + kk]))));
                }
            }
        }
#pragma unroll
    for (int q = 0; q < K * K; q++) res[q] = group_sum<G>(res[q]);
    matmul_m<K, D>(phi_inv, res, H);
}

// Local energy and O row of sample s from its state (sp, tl, phi) held by a group
// of G lanes: E_zz (rbm.cc:1012-1031), H_loc = Phi^-1 (diag(E_zz) Phi - h sum_i
// Phi^(i)) (rbm.cc:1042-1064) or the hs-like H_loc above, O_{k,p} = sum_kk c_{k,kk} d_{kk,p} with
// c_{k,kk} = (Phi^-1)_{k,kk} Phi_{kk,k} (Cramer's rule applied to rbm.cc:199-239).
template <bool CPLX, int K, int G>
__device__ inline void mx_measure(const MixedSampleArgs &a, const MxTab<CPLX> &tb, const double *Jm,
                                  const uint32_t *sp, const typename Ty<CPLX>::T *tl, int g,
                                  const typename Ty<CPLX>::D *phi, long long s, bool live)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const int N = a.N, M = a.M, NW = a.NW, L = a.L;
    const T *chT = tb.ch, *shT = tb.sh;
    const D *eaP = tb.eaP;
    const D *eaN = tb.eaN;
    D phi_inv[K * K];
    inv_m<K, D>(phi, phi_inv);

    D H[K * K];
    if (a.model != 0) {
        mx_hloc_hs<CPLX, K, G>(a, tb, Jm, sp, tl, g, phi, phi_inv, H);
    } else {
    // E_zz per replica: lanes take rows i = g, g+G, ..., and the columns j < i of row i
    // starting at a lane-dependent offset (rows of J share LDS banks)
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
    // sum_i Phi^(i): lanes take spins i = g, g + G, ... two at a time; each table entry
    // cosh/sinh(2 w_{kk,i,j}) read feeds all K replicas, so a lane runs 2 K independent
    // product chains.  The j loop starts at a lane-dependent offset to spread LDS banks.
    {
        D sx[K * K];
#pragma unroll
        for (int q = 0; q < K * K; q++) sx[q] = zeroD<D>();
#pragma unroll
        for (int kk = 0; kk < K; kk++) {
            for (int i0 = g; i0 < N; i0 += 2 * G) {
                const int i1 = (i0 + G < N) ? i0 + G : i0;  // i1 == i0: second chain discarded
                const bool two = i0 + G < N;
                int si0[K], si1[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    si0[k] = (sp[k * NW + (i0 >> 5)] >> (i0 & 31)) & 1u;
                    si1[k] = (sp[k * NW + (i1 >> 5)] >> (i1 & 31)) & 1u;
                }
                const T *ch0 = chT + ((long long)kk * N + i0) * M, *sh0 = shT + ((long long)kk * N + i0) * M;
                const T *ch1 = chT + ((long long)kk * N + i1) * M, *sh1 = shT + ((long long)kk * N + i1) * M;
                D prod0[K], prod1[K];
                T pf0[K], pf1[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    prod0[k] = oneD<D>();
                    prod1[k] = oneD<D>();
                    pf0[k] = oneT<T>();
                    pf1[k] = oneT<T>();
                }
                int j = g % M;
                for (int jj = 0; jj < M; jj++) {
                    const T c0 = ch0[j], s0 = sh0[j], c1 = ch1[j], s1 = sh1[j];
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        const T t = tl[(k * K + kk) * M + j];
                        pf0[k] = mulT(pf0[k], subT(c0, mulT(t, si0[k] ? negT(s0) : s0)));
                        pf1[k] = mulT(pf1[k], subT(c1, mulT(t, si1[k] ? negT(s1) : s1)));
                    }
                    if ((jj & 7) == 7) {
#pragma unroll
                        for (int k = 0; k < K; k++) {
                            prod0[k] = mulD(prod0[k], toD(pf0[k]));
                            prod1[k] = mulD(prod1[k], toD(pf1[k]));
                            pf0[k] = oneT<T>();
                            pf1[k] = oneT<T>();
                        }
                    }
                    if (++j == M) j = 0;
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    prod0[k] = mulD(prod0[k], toD(pf0[k]));
                    const D e0 = si0[k] ? eaP[kk * N + i0] : eaN[kk * N + i0];
                    sx[k * K + kk] = addD(sx[k * K + kk], mulD(mulD(phi[k * K + kk], e0), prod0[k]));
                    if (two) {
                        prod1[k] = mulD(prod1[k], toD(pf1[k]));
                        const D e1 = si1[k] ? eaP[kk * N + i1] : eaN[kk * N + i1];
                        sx[k * K + kk] = addD(sx[k * K + kk], mulD(mulD(phi[k * K + kk], e1), prod1[k]));
                    }
                }
            }
        }
        D res[K * K];
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                D sxs = group_sum<G>(sx[k * K + kk]);
                res[k * K + kk] = subD(scaleD(ez[k], phi[k * K + kk]), scaleD(a.h, sxs));
            }
        matmul_m<K, D>(phi_inv, res, H);
    }
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
                    ct[kk] = mulT(ck[kk], tl[(kk * K + k) * M + j]);
                    v = addT(v, ct[kk]);
                }
                if (Ub)
#pragma unroll
                    for (int kk = 0; kk < K; kk++) Ub[(long long)kk * a.Ucols + k * M + j] = ct[kk];
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
        if (Orow)
            for (long long q = a.P + g; q < a.Pld; q += G) Orow[q] = zeroT<T>();
    }
}

// Local energies and O rows of all S samples from their stored spins: G lanes per
// sample, 64/G samples per wavefront, each block looping over sample batches (grid
// sized to the CUs) so that the tables are staged in LDS once per block.  Sample
// parallelism (S) instead of chain parallelism (C) keeps every SIMD busy.
template <bool CPLX, int K, int G, int JMAX, bool TAB>
__global__ __launch_bounds__(256) void k_measure_mx(MixedSampleArgs a)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    constexpr int SPW = 64 / G;
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, M = a.M, NW = a.NW;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int g = lane & (G - 1), grp = lane / G;
    const int spb = (blockDim.x >> 6) * SPW;
    const int sib = wv * SPW + grp;
    const long long S = (long long)a.C * a.m;

    MxTab<CPLX> tb;
    unsigned char *p = mx_stage_tables<CPLX, TAB>(a, K, tb, smem);
    const double *Jm = a.J;
    if constexpr (TAB) {
        double *jl = (double *)p;
        for (int q = tid; q < N * N; q += blockDim.x) jl[q] = a.J[q];
        Jm = jl;
        p = (unsigned char *)(jl + N * N);
    }
    T *tl = (T *)p + (long long)sib * K * K * M;
    uint32_t *sp = (uint32_t *)((T *)p + (long long)spb * K * K * M) + sib * K * NW;
    int overflow = 0;
    for (long long base = (long long)blockIdx.x * spb; base < S; base += (long long)gridDim.x * spb) {
        const long long s_raw = base + sib;
        const bool live = s_raw < S;
        const long long s = live ? s_raw : S - 1;
        __syncthreads();  // previous batch done with sp / tl (and the tables staged)
        if (g == 0)
            for (int w = 0; w < K * NW; w++) sp[w] = a.spins[s * K * NW + w];
        __syncthreads();
        D phi[K * K];
        mx_refresh<CPLX, K, G, JMAX>(a, tb, sp, tl, g, 0, false, phi, overflow);
        mx_measure<CPLX, K, G>(a, tb, Jm, sp, tl, g, phi, s, live);
    }
    if (overflow) atomicOr(a.flags, 1u);
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

// out[col] = sum over chunks of part[chunk * stride + col], col < ncol, chunks in order
template <typename D>
__global__ void k_chunksum_mx(const D *__restrict__ part, long long ncol, long long stride, int nchunk,
                              D *__restrict__ out)
{
    long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    D acc = part[col];
    for (int ch = 1; ch < nchunk; ch++) acc = addD(acc, part[(long long)ch * stride + col]);
    out[col] = acc;
}

// partial sums over groups of CG consecutive chunks: grid (col tiles, groups)
template <typename D>
__global__ void k_chunkgroup_mx(const D *__restrict__ part, long long ncol, long long stride, int nchunk, int cg,
                                D *__restrict__ gout)
{
    long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    int c0 = blockIdx.y * cg, c1 = c0 + cg;
    if (c1 > nchunk) c1 = nchunk;
    D acc = zeroD<D>();
    for (int ch = c0; ch < c1; ch++) acc = addD(acc, part[(long long)ch * stride + col]);
    gout[(long long)blockIdx.y * ncol + col] = acc;
}

// out[col] = sum_chunks part[...], deterministic, parallel over column tiles and
// chunk groups; scratch holds ngroups * ncol values
template <typename D>
void chunk_reduce(const D *part, long long ncol, long long stride, int nchunk, D *scratch, D *out, hipStream_t st)
{
    const int cg = 32;
    int ngroups = (nchunk + cg - 1) / cg;
    if (ngroups <= 1) {
        hipLaunchKernelGGL(k_chunksum_mx<D>, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, st, part, ncol, stride,
                           nchunk, out);
        return;
    }
    hipLaunchKernelGGL(k_chunkgroup_mx<D>, dim3((unsigned)((ncol + 63) / 64), (unsigned)ngroups), dim3(64), 0, st,
                       part, ncol, stride, nchunk, cg, scratch);
    hipLaunchKernelGGL(k_chunksum_mx<D>, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, st, (const D *)scratch,
                       ncol, ncol, ngroups, out);
}

// ---------------------------------------------------------------- launchers
template <bool CPLX, int K, int G>
hipError_t launch_measure_g(const MixedSampleArgs &a, hipStream_t st, int wpb)
{
    using T = typename Ty<CPLX>::T;
    const int spb = wpb * (64 / G);
    size_t lds = (size_t)spb * K * K * a.M * sizeof(T) + (size_t)spb * K * a.NW * 4;
    const bool tab = a.tab_lds;  // tables and J in LDS
    if (tab) lds += mx_tab_bytes(CPLX, a.N, a.M, K) + (size_t)a.N * a.N * sizeof(double);
    const long long S = (long long)a.C * a.m;
    long long blocks = (S + spb - 1) / spb;
    const void *fn = tab ? (const void *)k_measure_mx<CPLX, K, G, 8, true> : (const void *)k_measure_mx<CPLX, K, G, 8, false>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    // resident blocks: limited by LDS (160 KB per CU) and by 8 waves per SIMD
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    long long per_cu = lds > 0 ? (long long)(160 * 1024) / (long long)lds : 8;
    if (per_cu > 32 / wpb) per_cu = 32 / wpb;
    if (per_cu < 1) per_cu = 1;
    const char *eb = getenv("NQES_MEASURE_BLOCKS_PER_CU");
    if (eb && *eb && atoi(eb) > 0) per_cu = atoi(eb);
    if (blocks > per_cu * ncu) blocks = per_cu * ncu;
    if (tab) hipLaunchKernelGGL((k_measure_mx<CPLX, K, G, 8, true>), dim3((unsigned)blocks), dim3(64 * wpb), lds, st, a);
    else hipLaunchKernelGGL((k_measure_mx<CPLX, K, G, 8, false>), dim3((unsigned)blocks), dim3(64 * wpb), lds, st, a);
    return hipGetLastError();
}

template <bool CPLX, int K>
hipError_t launch_measure_k(const MixedSampleArgs &a, hipStream_t st, int G, int wpb)
{
    switch (G) {
    case 16: return launch_measure_g<CPLX, K, 16>(a, st, wpb);
    case 32: return launch_measure_g<CPLX, K, 32>(a, st, wpb);
    case 64: return launch_measure_g<CPLX, K, 64>(a, st, wpb);
    default: return hipErrorInvalidValue;
    }
}

template <bool CPLX>
hipError_t launch_measure_c(int K, const MixedSampleArgs &a, hipStream_t st, int G, int wpb)
{
    switch (K) {
    case 1: return launch_measure_k<CPLX, 1>(a, st, G, wpb);
    case 2: return launch_measure_k<CPLX, 2>(a, st, G, wpb);
    case 3: return launch_measure_k<CPLX, 3>(a, st, G, wpb);
    case 4: return launch_measure_k<CPLX, 4>(a, st, G, wpb);
    default: return hipErrorInvalidValue;
    }
}

} // namespace

hipError_t nq_chunk_reduce(bool cplx, const void *part, long long ncol, long long stride, int nchunk,
                           void *scratch, void *out, hipStream_t st)
{
    if (cplx) chunk_reduce<cd>((const cd *)part, ncol, stride, nchunk, (cd *)scratch, (cd *)out, st);
    else chunk_reduce<double>((const double *)part, ncol, stride, nchunk, (double *)scratch, (double *)out, st);
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
    dim3 grid((unsigned)((rows + 15) / 16), (unsigned)((ncol + 15) / 16));
    hipLaunchKernelGGL(k_theta_gemm, grid, dim3(64), 0, st, N, M, K, NW, R, rows, spins0, (const float *)wT,
                       (const float *)bT, (float *)theta0);
    return hipGetLastError();
}

hipError_t nq_launch_measure_mx(bool cplx, int K, const MixedSampleArgs &a, int G, int wpb, hipStream_t st)
{
    return cplx ? launch_measure_c<true>(K, a, st, G, wpb) : launch_measure_c<false>(K, a, st, G, wpb);
}

size_t nq_sample_mx_lds(bool cplx, int N, int M, int K, int G, int wpb, bool tab_lds)
{
    size_t ts = cplx ? 8 : 4;
    int NW = (N + 31) / 32;
    int cpb = wpb * (64 / G);
    size_t lds = (size_t)cpb * K * K * M * ts + (size_t)cpb * K * NW * 4;
    if (tab_lds) lds += mx_tab_bytes(cplx, N, M, K);
    return lds;
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
        hipError_t e = hipFuncSetAttribute((const void *)k_sv_rows_mx<CPLX, QV, WPR, NP>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
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
