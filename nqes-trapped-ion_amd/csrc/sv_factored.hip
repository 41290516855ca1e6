// sv_factored.hip -- S*v of the mixed path from per-sample factors instead of O.
//
// With Cramer's rule, the O row of sample s is, for RBM k (rbm.cc:199-239),
//   a block  O_{k,i}   = sum_kk c_{s,k,kk} sigma^kk_i
//   b block  O_{k,j}   = sum_kk c_{s,k,kk} t_{s,kk,k,j}
//   w block  O_{k,i,j} = sum_kk c_{s,k,kk} sigma^kk_i t_{s,kk,k,j}
// where sigma^kk = +-1 are replica kk's spins and t_{s,kk,k,j} = tanh theta_j of RBM k
// at those spins.  Rows r = s K + kk of the factor matrix U hold
//   U[r][k M + j] = c_{s,k,kk} t_{s,kk,k,j},  U[r][K M + k] = c_{s,k,kk},
// zero-padded to Ucols = 16 ceil((K M + K) / 16) columns.  With B[i][k M + j] =
// v_w[k][i][j], B[i][K M + k] = v_a[k][i] and bias[k M + j] = v_b[k][j]:
//   O_s^T v = sum_kk sum_col U[r][col] (sigma_r^T B[:, col] + bias[col])
//   sum_s conj(O_s) w_s:  a and w blocks  G[i][col] = sum_r sigma_r[i] conj(U[r][col]) w_s(r)
//                         b block         sum_r conj(U[r][col]) w_s(r)
// which is the reference's S*v (rbm.cc:661-684, matrix-free: the per-sample state is
// stored instead of O) as two GEMMs with the +-1 spin matrix as one operand.  Both run
// on FP64 MFMA (v_mfma_f64_16x16x4f64): FP64 arithmetic on FP32 factors, so the
// operator is linear to FP64 rounding, as the stored-O pass is.
//
// Pass 1 k_svf_fwd     one wavefront per 16 rows: Y = Sigma B on MFMA, then
//                      w_r = sum_col U (Y + bias) -> wrow[r]
// Pass 2 k_svf_combine w_s = sum_kk wrow[s K + kk] - o^T v
// Pass 3 k_svf_bwd     blocks over row chunks, wavefronts over column tiles (each with
//                      all its i tiles): G on MFMA, b-block column sums, sum_s w_s; one
//                      partial raw vector (P + 1) per block, summed in block order after.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>
#include <stdlib.h>

#include "cplx64.hpp"
#include "kernels.hpp"
#include "mixed_math.hpp"

using namespace nqm;
using nq64::cd;

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline double reD(double x) { return x; }
__device__ inline double reD(cd z) { return z.re; }
__device__ inline double imD(double) { return 0.0; }
__device__ inline double imD(cd z) { return z.im; }
template <typename D> __device__ inline D mkD(double re, double im);
template <> __device__ inline double mkD<double>(double re, double) { return re; }
template <> __device__ inline cd mkD<cd>(double re, double im) { return nq64::mk(re, im); }

// +-1 of spin i (bit set: sigma = -1) from the row's two spin words (N <= 64)
__device__ inline double spin_pm(uint32_t w0, uint32_t w1, int i)
{
    const uint32_t w = (i < 32) ? w0 : w1;
    return ((w >> (i & 31)) & 1u) ? -1.0 : 1.0;
}

// columns [c0, c0 + W) of B (N4 x W) and of bias (W) from v, in LDS
template <typename D>
__device__ inline void stage_B(const SvfArgs &a, D *Bl, D *bias, int c0, int W)
{
    const D *v = (const D *)a.v;
    const int N = a.N, M = a.M, K = a.K;
    const int N4 = a.n4;
    const long long KM = (long long)K * M;
    for (long long q = threadIdx.x; q < (long long)N4 * W; q += blockDim.x) {
        const int i = (int)(q / W), col = c0 + (int)(q % W);
        D b = zeroD<D>();
        if (i < N) {
            if (col < KM) {
                const int k = col / M, j = col % M;
                b = v[(long long)k * a.L + N + M + (long long)i * M + j];
            } else if (col < KM + K) {
                b = v[(long long)(col - KM) * a.L + i];
            }
        }
        Bl[q] = b;
    }
    for (int q = threadIdx.x; q < W; q += blockDim.x) {
        const int col = c0 + q;
        D b = zeroD<D>();
        if (col < KM) b = v[(long long)(col / M) * a.L + N + col % M];
        bias[q] = b;
    }
}

// v_mfma_f64_16x16x4f64 operand layout: lane l holds A[l%16][l/16] and B[l/16][l%16];
// result register e of lane l is D[4 e + l/16][l%16] (rows interleaved by lane group,
// unlike the FP32 16x16x4 form whose register e holds row 4 (l/16) + e).
//
// Pass 1.  Block (b, g) handles column group g (tiles [g cgt, min((g+1) cgt, NCT)),
// its part of B staged in LDS) and loops with its 4 wavefronts over 16-row tiles.
// Lane l: A operand row r0 + l%16 (spins i = 4 q + l/16), result rows r0 + 4 e + l/16
// (e < 4) at column l%16 of each 16-column tile.  NS = N4/4 steps of the K dimension
// (N <= 64).  Output: the group's partial row sums wrow[g R + r].
template <bool CPLX, int NSMAX>
__global__ __launch_bounds__(256) void k_svf_fwd(SvfArgs a, int cgt)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    extern __shared__ __align__(16) unsigned char smem_f[];
    const int N = a.N, K = a.K, NW = a.NW, UC = a.Ucols;
    const int N4 = a.n4;
    const int NS = N4 / 4;
    const int NCT = UC / 16;
    const int ct0 = blockIdx.y * cgt;
    const int nct = (NCT - ct0 < cgt) ? NCT - ct0 : cgt;
    const int W = 16 * nct;
    D *Bl = (D *)smem_f;
    D *bias = Bl + (long long)N4 * W;
    stage_B<D>(a, Bl, bias, 16 * ct0, W);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int hi = lane >> 4, lo = lane & 15;
    const long long R = a.S * K;
    const long long ntile = (R + 15) / 16;
    const T *U = (const T *)a.U;
    D *wrow = (D *)a.wrow + (long long)blockIdx.y * R;
    for (long long t = (long long)blockIdx.x * 4 + wv; t < ntile; t += (long long)gridDim.x * 4) {
        const long long r0 = t * 16;
        const long long ra = r0 + lo;
        uint32_t w0 = 0, w1 = 0;
        if (ra < R) {
            w0 = a.spins[ra * NW];
            if (NW > 1) w1 = a.spins[ra * NW + 1];
        }
        double av[NSMAX];
#pragma unroll
        for (int q = 0; q < NSMAX; q++) {
            const int i = 4 * q + hi;
            av[q] = (q < NS && ra < R && i < N) ? spin_pm(w0, w1, i) : 0.0;
        }
        D rs[4];
#pragma unroll
        for (int q = 0; q < 4; q++) rs[q] = zeroD<D>();
        for (int c = 0; c < nct; c++) {
            const int cl = c * 16 + lo, col = 16 * ct0 + cl;
            v4d accR = {0.0, 0.0, 0.0, 0.0};
            v4d accI = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < NSMAX; q++) {
                if (q < NS) {
                    const D bv = Bl[(long long)(4 * q + hi) * W + cl];
                    accR = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], reD(bv), accR, 0, 0, 0);
                    if constexpr (CPLX) accI = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], imD(bv), accI, 0, 0, 0);
                }
            }
            const D bc = bias[cl];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long long r = r0 + 4 * q + hi;
                if (r < R) {
                    const D z = addD(mkD<D>(accR[q], accI[q]), bc);
                    rs[q] = addD(rs[q], mulD(toD(U[r * UC + col]), z));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) rs[q] = addD(rs[q], shflxD(rs[q], off));
        if (lo == 0)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long long r = r0 + 4 * q + hi;
                if (r < R) wrow[r] = rs[q];
            }
    }
}

// Pass 2.  w_s = sum_g sum_kk wrow[g R + s K + kk] - o^T v
template <typename D>
__global__ void k_svf_combine(SvfArgs a)
{
    const D *wrow = (const D *)a.wrow;
    D *w = (D *)a.w;
    const D ov = *(const D *)a.ov;
    const long long R = a.S * a.K;
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < a.S;
         s += (long long)gridDim.x * blockDim.x) {
        D acc = zeroD<D>();
        for (int g = 0; g < a.ncg; g++)
            for (int kk = 0; kk < a.K; kk++) acc = addD(acc, wrow[g * R + s * a.K + kk]);
        w[s] = subD(acc, ov);
    }
}

// Pass 3.  Block (b, g) takes the 4-row steps [b n4 / nblk, (b+1) n4 / nblk) and the
// column tiles of group g: wavefront wv owns tile ct = g GS + wv (GS wavefronts per
// block) with all its i tiles it < NIT = ceil(N / 16), so each U value is loaded once
// and feeds NIT MFMAs.  Lane l: A operand sigma_{r0 + l/16}[16 it + l%16], B operand
// conj(U[r0 + l/16][16 ct + l%16]) w_s; result rows i = 16 it + 4 e + l/16 (e < 4) at
// column 16 ct + l%16.  The B operand is also summed over rows (b block), and group 0's
// wavefront 0 sums w_s over rows with kk = 0.  The groups of block row b write disjoint
// entries of part[b][0..P].
template <bool CPLX, int NITM, int WSEL>
__global__ __launch_bounds__(1024) void k_svf_bwd(SvfArgs a)
{
    using T = typename Ty<CPLX>::T;
    using D = typename Ty<CPLX>::D;
    const int N = a.N, M = a.M, K = a.K, NW = a.NW, UC = a.Ucols;
    const int NCT = UC / 16, NIT = (N + 15) / 16;
    const int GS = blockDim.x >> 6;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int hi = lane >> 4, lo = lane & 15;
    const int ct = blockIdx.y * GS + wv;
    const bool tile = ct < NCT;
    const bool wsum = blockIdx.y == 0 && wv == 0;
    const long long R = a.S * K;
    const long long KM = (long long)K * M;
    const long long n4 = (R + 3) / 4;
    const long long q0 = n4 * blockIdx.x / gridDim.x, q1 = n4 * (blockIdx.x + 1) / gridDim.x;
    const T *U = (const T *)a.U;
    const D *w = (const D *)a.w;

    v4d accR[NITM], accI[NITM];
#pragma unroll
    for (int it = 0; it < NITM; it++) {
        accR[it] = v4d{0.0, 0.0, 0.0, 0.0};
        accI[it] = v4d{0.0, 0.0, 0.0, 0.0};
    }
    D cs = zeroD<D>(), ws = zeroD<D>();
    if (tile || wsum)
        for (long long step = q0; step < q1; step++) {
            const long long r = step * 4 + hi;  // this lane's row (K dimension)
            const bool ok = r < R;
            uint32_t w0 = 0, w1 = 0;
            D wr = zeroD<D>();
            T uv = zeroT<T>();
            if (ok) {
                w0 = a.spins[r * NW];
                if (NW > 1) w1 = a.spins[r * NW + 1];
                if constexpr (WSEL == 0) wr = w[r / K];
                else if constexpr (WSEL == 1) wr = mkD<D>(1.0, 0.0);
                else {
                    const cd e = ((const cd *)a.E)[r / K];
                    wr = mkD<D>(e.re, e.im);
                }
                if (tile) uv = U[r * UC + 16 * ct + lo];
                if (wsum && lo == 0 && (r % K) == 0) ws = addD(ws, wr);
            }
            if (tile) {
                const D bv = mulD(conjD(toD(uv)), wr);
                cs = addD(cs, bv);
#pragma unroll
                for (int it = 0; it < NITM; it++) {
                    if (it < NIT) {
                        const int i = 16 * it + lo;
                        const double av = (ok && i < N) ? spin_pm(w0, w1, i) : 0.0;
                        accR[it] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, reD(bv), accR[it], 0, 0, 0);
                        if constexpr (CPLX)
                            accI[it] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, imD(bv), accI[it], 0, 0, 0);
                    }
                }
            }
        }
    D *part = (D *)a.part + (long long)blockIdx.x * (a.P + 1);
    if (tile) {
        const int col = 16 * ct + lo;
#pragma unroll
        for (int it = 0; it < NITM; it++) {
            if (it < NIT) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int i = 16 * it + 4 * e + hi;
                    if (i < N) {
                        const D g = mkD<D>(accR[it][e], accI[it][e]);
                        if (col < KM) part[(col / M) * a.L + N + M + (long long)i * M + col % M] = g;
                        else if (col < KM + K) part[(col - KM) * a.L + i] = g;
                    }
                }
            }
        }
        D t = cs;
        t = addD(t, shflxD(t, 16));
        t = addD(t, shflxD(t, 32));
        if (hi == 0 && col < KM) part[(col / M) * a.L + N + col % M] = t;
    }
    if (wsum) {
        D t = ws;
        t = addD(t, shflxD(t, 16));
        t = addD(t, shflxD(t, 32));
        if (lane == 0) part[a.P] = t;
    }
}

// backward pass with weights w (WSEL 0), ones (1) or the local energies (2)
template <bool CPLX, int WSEL>
hipError_t launch_bwd(const SvfArgs &a, const SvfShape &sh, hipStream_t st)
{
    const dim3 grid(sh.bwd_blocks, sh.bwd_cgs), bb(64 * sh.bwd_waves);
    if (a.N <= 32) hipLaunchKernelGGL((k_svf_bwd<CPLX, 2, WSEL>), grid, bb, 0, st, a);
    else hipLaunchKernelGGL((k_svf_bwd<CPLX, 4, WSEL>), grid, bb, 0, st, a);
    return hipGetLastError();
}

template <bool CPLX>
hipError_t launch_svf(const SvfArgs &a0, const SvfShape &sh, void *scratch, void *out, hipStream_t st)
{
    using D = typename Ty<CPLX>::D;
    SvfArgs a = a0;
    a.ncg = sh.fwd_cgs;
    const int W = 16 * sh.fwd_cgt;
    const size_t lds = ((size_t)a.n4 * W + W) * sizeof(D);
    if (lds > 65536) return hipErrorInvalidValue;  // groups are sized to fit
    const dim3 grid(sh.fwd_blocks, sh.fwd_cgs);
    if (a.n4 <= 32)
        hipLaunchKernelGGL((k_svf_fwd<CPLX, 8>), grid, dim3(256), lds, st, a, sh.fwd_cgt);
    else
        hipLaunchKernelGGL((k_svf_fwd<CPLX, 16>), grid, dim3(256), lds, st, a, sh.fwd_cgt);
    hipLaunchKernelGGL(k_svf_combine<D>, dim3(1024), dim3(256), 0, st, a);
    hipError_t eb = launch_bwd<CPLX, 0>(a, sh, st);
    if (eb != hipSuccess) return eb;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    nq_chunk_reduce(CPLX, a.part, a.P + 1, a.P + 1, sh.bwd_blocks, scratch, out, st);
    return hipGetLastError();
}

} // namespace

// Sum over samples of |O_sp|^2 from the factors.  With U_kk = U[s K + kk][col]:
//   a and w blocks  sum_s sum_kk |U_kk|^2 + 2 sum_{kk < kk'} sum_s tau_s[i] Re(U_kk conj U_kk'),
//                   tau_s[i] = sigma^kk_i sigma^kk'_i (+-1),
//   b block         sum_s |sum_kk U_kk|^2.
// Grid (blocks over sample chunks, replica pairs); wavefronts over column tiles.  The
// cross term of pair (kk, kk') is a GEMM [N x S] [S x Ucols] on FP64 MFMA (A operand tau,
// B operand Re(U_kk conj U_kk')).  Pair 0 also adds the diagonal and b-block sums.
// Block (b, pair) writes a full partial vector part[pair nblk + b][0..P).
template <bool CPLX, int NITM, int K>
__global__ __launch_bounds__(1024) void k_svf_sq(SvfArgs a)
{
    using T = typename Ty<CPLX>::T;
    constexpr int NW = NITM / 2;  // spin words: 1 for up to 32 spins, 2 for up to 64
    const int N = a.N, M = a.M, UC = a.Ucols;
    const int NCT = UC / 16, NIT = (N + 15) / 16;
    const int NWV = blockDim.x >> 6;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int hi = lane >> 4, lo = lane & 15;
    const long long S = a.S, KM = (long long)K * M;
    const long long n4 = (S + 3) / 4;
    const long long q0 = n4 * blockIdx.x / gridDim.x, q1 = n4 * (blockIdx.x + 1) / gridDim.x;
    // replica pair of this block: pairs enumerated (0,1), (0,2), .., (1,2), ..; K = 1: none
    const int pair = blockIdx.y;
    int ka = 0, kb = 0;
    {
        int q = 0;
        for (int x = 0; x < K; x++)
            for (int y = x + 1; y < K; y++, q++)
                if (q == pair) {
                    ka = x;
                    kb = y;
                }
    }
    constexpr bool cross = K > 1;
    const T *U = (const T *)a.U;
    double *part = (double *)a.part + ((long long)pair * gridDim.x + blockIdx.x) * (a.P + 1);
    for (int ct = wv; ct < NCT; ct += NWV) {
        const int col = 16 * ct + lo;
        v4d acc[NITM];
#pragma unroll
        for (int it = 0; it < NITM; it++) acc[it] = v4d{0.0, 0.0, 0.0, 0.0};
        double dg = 0.0, bq = 0.0;
        // loads unconditional (sample index clamped; a step past S adds exact zeros), K,
        // the spin words and the pair-0 terms fixed at compile time, MFMAs of spin tiles
        // past N fed zeros: a step is straight-line code, and four steps issue their loads
        // before their sums, which run in the original order
        // a step: its loads (Step) and its sums and MFMAs (use)
        struct Step {
            bool ok;
            T u[K];
            uint32_t wa0, wb0, wa1, wb1;  // spin words of replicas kk, kk' (tau: their xor)
        };
        auto load = [&](long long step, Step &x) __attribute__((always_inline)) {
            const long long s0 = step * 4 + hi;
            x.ok = s0 < S;
            const long long s = x.ok ? s0 : S - 1;
            const T *Us = U + s * K * UC + col;
#pragma unroll
            for (int kk = 0; kk < K; kk++) x.u[kk] = Us[kk * UC];
            x.wa0 = x.wb0 = x.wa1 = x.wb1 = 0;
            if constexpr (cross) {
                x.wa0 = a.spins[(s * K + ka) * NW];
                x.wb0 = a.spins[(s * K + kb) * NW];
                if constexpr (NW > 1) {
                    x.wa1 = a.spins[(s * K + ka) * NW + 1];
                    x.wb1 = a.spins[(s * K + kb) * NW + 1];
                }
            }
        };
        auto steps = [&](auto p0c) {
            constexpr bool P0 = decltype(p0c)::value;
            auto use = [&](const Step &x) __attribute__((always_inline)) {
                const bool ok = x.ok;
                if constexpr (P0) {
                    double br = 0.0, bi = 0.0;
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        const T u = x.u[kk];
                        if constexpr (CPLX) {
                            dg += ok ? (double)u.re * u.re + (double)u.im * u.im : 0.0;
                            br += u.re;
                            bi += u.im;
                        } else {
                            dg += ok ? (double)u * u : 0.0;
                            br += u;
                        }
                    }
                    bq += ok ? br * br + bi * bi : 0.0;
                }
                if constexpr (cross) {
                    // ka, kb are runtime: pick the pair's two factors
                    T ua = x.u[0], ub = x.u[0];
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        if (kk == ka) ua = x.u[kk];
                        if (kk == kb) ub = x.u[kk];
                    }
                    double vab;
                    if constexpr (CPLX) vab = (double)ua.re * ub.re + (double)ua.im * ub.im;
                    else vab = (double)ua * ub;
                    vab = ok ? vab : 0.0;
#pragma unroll
                    for (int it = 0; it < NITM; it++) {
                        const int i = 16 * it + lo;
                        const double av = (ok && i < N) ? spin_pm(x.wa0 ^ x.wb0, x.wa1 ^ x.wb1, i) : 0.0;
                        acc[it] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, vab, acc[it], 0, 0, 0);
                    }
                }
            };
            // unrolled by hand (the compiler does not unroll a loop of MFMAs with a
            // runtime trip count): four steps' loads, then their sums in order
            long long step = q0;
            for (; step + 4 <= q1; step += 4) {
                Step x[4];
#pragma unroll
                for (int u = 0; u < 4; u++) load(step + u, x[u]);
#pragma unroll
                for (int u = 0; u < 4; u++) use(x[u]);
            }
            for (; step < q1; step++) {
                Step x;
                load(step, x);
                use(x);
            }
        };
        if (pair == 0) steps(std::true_type{});
        else steps(std::false_type{});
        // column sums over the 4 sample lanes of each column
        dg += __shfl_xor(dg, 16);
        dg += __shfl_xor(dg, 32);
        bq += __shfl_xor(bq, 16);
        bq += __shfl_xor(bq, 32);
        if (col < KM + K) {
            const double dcol = dg;  // every lane of column col holds its diagonal sum
#pragma unroll
            for (int it = 0; it < NITM; it++) {
                if (it < NIT) {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int i = 16 * it + 4 * e + hi;
                        if (i < N) {
                            const double g = 2.0 * acc[it][e] + (pair == 0 ? dcol : 0.0);
                            if (col < KM) part[(col / M) * a.L + N + M + (long long)i * M + col % M] = g;
                            else part[(col - KM) * a.L + i] = g;
                        }
                    }
                }
            }
            if (hi == 0 && col < KM) part[(col / M) * a.L + N + col % M] = (pair == 0) ? bq : 0.0;
        }
    }
}

// so = conj(sum_s conj O_s) (w = 1), sf = sum_s E conj O_s, ss = sum |O|^2 -> complex arrays
template <typename D>
__global__ void k_svf_pack_moments(long long P, const D *r1, const D *rE, const double *rsq, cd *so, cd *ss, cd *sf)
{
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long long)gridDim.x * blockDim.x) {
        const cd x1 = nq64::mk(reD(r1[p]), imD(r1[p])), xe = nq64::mk(reD(rE[p]), imD(rE[p]));
        so[p] = nq64::conj(x1);
        sf[p] = xe;
        ss[p] = nq64::mk(rsq[p], 0.0);
    }
}

// sum |O|^2 into rsq = tmp + 2 (P + 1) D, then the three moment vectors from r1 (w = 1,
// at tmp) and rE (w = E, at tmp + (P + 1) D)
template <bool CPLX>
hipError_t launch_sq_pack(const SvfArgs &a0, const SvfShape &sh, void *part, void *scratch, void *tmp, void *so,
                          void *ss, void *sf, hipStream_t st)
{
    using D = typename Ty<CPLX>::D;
    const long long P = a0.P;
    D *r1 = (D *)tmp, *rE = r1 + (P + 1);
    double *rsq = (double *)(rE + (P + 1));
    SvfArgs a = a0;
    a.part = part;
    const int npair = a0.K > 1 ? a0.K * (a0.K - 1) / 2 : 1;
    const int nct = a0.Ucols / 16;
    const int waves = nct < 16 ? nct : 16;
    const dim3 grid(sh.bwd_blocks, npair), blk(64 * waves);
#define NQ_SQ(NITM, KK) hipLaunchKernelGGL((k_svf_sq<CPLX, NITM, KK>), grid, blk, 0, st, a)
    if (a0.NW != (a0.N <= 32 ? 1 : 2)) return hipErrorInvalidValue;
    switch ((a0.N <= 32 ? 0 : 8) + a0.K) {
    case 1: NQ_SQ(2, 1); break;
    case 2: NQ_SQ(2, 2); break;
    case 3: NQ_SQ(2, 3); break;
    case 4: NQ_SQ(2, 4); break;
    case 9: NQ_SQ(4, 1); break;
    case 10: NQ_SQ(4, 2); break;
    case 11: NQ_SQ(4, 3); break;
    case 12: NQ_SQ(4, 4); break;
    default: return hipErrorInvalidValue;
    }
#undef NQ_SQ
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    nq_chunk_reduce(false, part, P, P + 1, sh.bwd_blocks * npair, scratch, rsq, st);
    hipLaunchKernelGGL(k_svf_pack_moments<D>, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, P, r1, rE, rsq,
                       (cd *)so, (cd *)ss, (cd *)sf);
    return hipGetLastError();
}

template <bool CPLX>
hipError_t launch_svf_moments(const SvfArgs &a0, const SvfShape &sh, void *part, void *scratch, void *tmp, void *so,
                              void *ss, void *sf, hipStream_t st)
{
    using D = typename Ty<CPLX>::D;
    const long long P = a0.P;
    D *r1 = (D *)tmp, *rE = r1 + (P + 1);
    double *rsq = (double *)(rE + (P + 1));
    const bool n32 = a0.N <= 32;
    SvfArgs a = a0;
    a.part = part;
    for (int pass = 0; pass < 2; pass++) {
        hipError_t e = pass == 0 ? launch_bwd<CPLX, 1>(a, sh, st) : launch_bwd<CPLX, 2>(a, sh, st);
        if (e != hipSuccess) return e;
        nq_chunk_reduce(CPLX, part, P + 1, P + 1, sh.bwd_blocks, scratch, pass == 0 ? (void *)r1 : (void *)rE, st);
    }
    return launch_sq_pack<CPLX>(a0, sh, part, scratch, tmp, so, ss, sf, st);
}

// Factored S*v applies when N <= 64 (two spin words per row, four i tiles).  Forward
// column groups: as many 16-column tiles as keep B (N4 x 16 cgt in D) within 64 KB of
// LDS.  Backward column groups: 16 tiles (one wavefront each) per block.
SvfShape nq_svf_shape(bool cplx, int N, int M, int K)
{
    SvfShape sh{};
    if (N > 64) return sh;
    const int UC = 16 * ((K * M + K + 15) / 16);
    const int nct = UC / 16;
    const int N4 = (N + 3) & ~3;
    const size_t dsz = cplx ? 16 : 8;
    int cgt = (int)(65536 / ((size_t)(N4 + 1) * 16 * dsz));
    if (cgt > nct) cgt = nct;
    if (cgt < 1) return sh;
    sh.ucols = UC;
    sh.ok = 1;
    sh.fwd_cgt = cgt;
    sh.fwd_cgs = (nct + cgt - 1) / cgt;
    sh.bwd_waves = nct < 16 ? nct : 16;
    sh.bwd_cgs = (nct + 15) / 16;
    sh.bwd_blocks = 2048;
    sh.fwd_blocks = 1024;
    const char *e = getenv("NQES_SVF_BWD_BLOCKS");
    if (e && *e && atoi(e) > 0) sh.bwd_blocks = atoi(e);
    e = getenv("NQES_SVF_FWD_BLOCKS");
    if (e && *e && atoi(e) > 0) sh.fwd_blocks = atoi(e);
    return sh;
}

hipError_t nq_launch_svf_moments(bool cplx, const SvfArgs &a, const SvfShape &sh, void *part, void *scratch, void *tmp,
                                 void *so, void *ss, void *sf, hipStream_t st)
{
    return cplx ? launch_svf_moments<true>(a, sh, part, scratch, tmp, so, ss, sf, st)
                : launch_svf_moments<false>(a, sh, part, scratch, tmp, so, ss, sf, st);
}

hipError_t nq_launch_svf_sq_pack(bool cplx, const SvfArgs &a, const SvfShape &sh, void *part, void *scratch,
                                 void *tmp, void *so, void *ss, void *sf, hipStream_t st)
{
    return cplx ? launch_sq_pack<true>(a, sh, part, scratch, tmp, so, ss, sf, st)
                : launch_sq_pack<false>(a, sh, part, scratch, tmp, so, ss, sf, st);
}

hipError_t nq_launch_svf_combine(bool cplx, const SvfArgs &a, hipStream_t st)
{
    if (cplx) hipLaunchKernelGGL(k_svf_combine<cd>, dim3(1024), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_svf_combine<double>, dim3(1024), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t nq_launch_svf(bool cplx, const SvfArgs &a, const SvfShape &sh, void *scratch, void *out, hipStream_t st)
{
    return cplx ? launch_svf<true>(a, sh, scratch, out, st) : launch_svf<false>(a, sh, scratch, out, st);
}
