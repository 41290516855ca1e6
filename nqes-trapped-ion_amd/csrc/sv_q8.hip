// sv_q8.hip -- S*v of the mixed path (real weights) from the factors on int8 MFMA.
//
// The factored S*v (rbm.cc:661-684 with O_s written as sum_kk c sigma^kk t, see
// sv_factored.hip) is two GEMMs with the +-1 spin matrix as one operand.  Here both
// run on v_mfma_i32_16x16x{32,64}_i8, whose integer results are exact.  Numbers are
// split into balanced base-256 digits (int8, least significant first):
//   factors   per row r a scale f_r = XMAX / 2^e_r with 2^e_r > max_col |U[r][col]|
//             and X = rint(U f_r) (|X| <= XMAX = 8355711, the largest number three
//             balanced digits hold; 23 bits, finer than the FP32 factors' 24-bit
//             mantissa relative to the row maximum), three digits -- U~ = X / f_r
//             differs from U by at most 2^-24 of the row's largest entry; done once per
//             SR step;
//   v         per column col of B (spins i x columns) a power of two 2^g > max_i |B|,
//             Q = rint(B 2^(46 - g)), six digits;
//   weights   per backward block a power of two 2^G > max |w_s(r) / f_r|,
//             Omega = rint(w / f_r 2^(46 - G)), six digits.
// The forward pass forms Y = Sigma Q (six MFMAs per tile, exact) and
//   w_r = (1 / f_r) sum_col X (Y 2^(g - 46) + bias) in FP64.
// For N <= 32 and at most 256 columns the forward pass is instead one GEMM over the
// columns with one scale for all of v (sv_q8_gemm.hip, on a row-major copy Uf of the
// factor digits that the pack kernel writes next to the tiles); NQES_Q8_FWD=tiles keeps
// the tiled form.
// The backward pass writes sigma = 1 - 2 b (b = 1 for spin down) and forms
//   sum_r sigma_r[i] X_r[col] Omega_r = T[col] - 2 sum_r b_r[i] X_r[col] Omega_r,
// the second sum digit by digit on MFMA (A operand = weight digits masked by b, 18
// MFMAs per tile pair), T (also the b-block column sums) by the same MFMAs on the
// unmasked weight digits (18 per column tile and chunk); products
// of equal digit weight 256^s share an int32 accumulator (exact), which goes to FP64
// once per block.  Both passes leave out the digit pairs of weight 256^0 and 256^1
// (QSKIP), which add at most 2^-46 of the largest product, so 15 of the 18 MFMAs of a
// digit set run.  The operator is the product of U~ with v and w quantised to 2^-47 of
// their scale (w per block of rows), to that truncation: linear and symmetric to ~1e-13.
//
// Layout of the quantised factors Uq: 16 x 16 tiles (row tile rt, column tile ct) of
// 768 bytes, tile (rt, ct) at dword (rt NCT + ct) 192; lane l of a tile holds dwords
// 3 l + t (digit plane t), whose byte j is digit t of row 16 rt + 4 (l/16) + j at
// column 16 ct + l%16.  For the forward pass that is the result layout of the 16x16
// MFMA (register j of lane l = row 4 (l/16) + j); for the backward pass dword t of row
// tile q is byte group q of the B operand of digit t of the 16x16x64 MFMA over four row
// tiles (lane l byte 4 q + j <-> row 16 q + 4 (l/16) + j).  Rows are padded to a
// multiple of 64 with zeros.
//
// MFMA int8 layouts (measured): A operand lane l byte e = A[l%16][k(l, e)], B operand
// lane l byte e = B[k(l, e)][l%16] for the same k(l, e) (any bijection onto the K
// dimension), result register e of lane l = D[4 (l/16) + e][l%16].
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cplx64.hpp"
#include "kernels.hpp"
#include "sv_q8_common.hpp"

namespace {

// four 0/1 bits -> four 0/1 bytes
__device__ inline uint32_t expand4(uint32_t b) { return ((b & 15u) * 0x00204081u) & 0x01010101u; }

// X of element j of a lane's three digit-plane dwords
__device__ inline int decode_x(uint32_t d0, uint32_t d1, uint32_t d2, int j)
{
    const int s = 8 * j;
    return (((int)(d2 << (24 - s)) >> 24) << 16) + (((int)(d1 << (24 - s)) >> 24) << 8) + ((int)(d0 << (24 - s)) >> 24);
}

__device__ inline double xor_sum16(double x)
{
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) x += __shfl_xor(x, off);
    return x;
}

// Pack: U (FP32, [R][UC] row-major) -> Uq tiles (and, when Uf is given, the same digits
// row-major for the GEMM forward pass) and escale[r] = 1 / f_r, the scale from the row
// maxima Umax when given (else a first pass over U).  One wavefront per 16-row tile,
// rows [R, Rp) written as zeros.  Complex factor rows (CM = 1, 2, 3: U holds (re, im)
// float pairs, Umax the row maxima of |re| + |im|): the digits of X_R = rint(re f_r),
// X_I = rint(im f_r) or X_A = X_R + X_I, with f_r = (XMAX - 1) / 2^e_r so that
// |X_A| <= XMAX.  QC < UC (c columns apart): digits of the first QC columns only, and the
// integers X of the K c columns (columns KM .. KM + K - 1) into Xc[r K + k] (exact floats:
// |X| <= XMAX < 2^24).
template <int CM>
__global__ __launch_bounds__(256) void k_q8_pack(const float *U, long long R, long long ntile, int UC, uint32_t *Uq,
                                                 double *escale, uint4 *Uf, const float *Umax, int QC, float *Xc,
                                                 int KM, int K)
{
    __shared__ __align__(16) uint8_t tile_t[4][QU * 256];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, lo = lane & 15;
    const int NCT = UC / 16, NQT = QC / 16;
    for (long long rt = (long long)blockIdx.x * 4 + wv; rt < ntile; rt += (long long)gridDim.x * 4) {
        float mx[4] = {0.f, 0.f, 0.f, 0.f};
        if (Umax) {  // the row maxima the measurement kernel wrote with the rows
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long long r = rt * 16 + 4 * g + j;
                if (r < R) mx[j] = Umax[r];
            }
        } else {
            for (int ct = 0; ct < NCT; ct++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const long long r = rt * 16 + 4 * g + j;
                    if (r < R) {
                        if constexpr (CM == 0) {
                            mx[j] = fmaxf(mx[j], fabsf(U[r * UC + 16 * ct + lo]));
                        } else {
                            const float2 u = ((const float2 *)U)[r * UC + 16 * ct + lo];
                            mx[j] = fmaxf(mx[j], fabsf(u.x) + fabsf(u.y));
                        }
                    }
                }
        }
        double sc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) mx[j] = fmaxf(mx[j], __shfl_xor(mx[j], off));
            int e = 0;
            if (mx[j] > 0.f) (void)frexpf(mx[j], &e);  // mx < 2^e
            sc[j] = ldexp(CM ? XMAX - 1.0 : XMAX, -e);
            const long long r = rt * 16 + 4 * g + j;
            if (lo == 0) escale[r] = (r < R && mx[j] > 0.f) ? 1.0 / sc[j] : 0.0;
        }
        if (Xc && lo < K) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long long r = rt * 16 + 4 * g + j;
                long long x;
                if constexpr (CM == 0) {
                    const float u = (r < R) ? U[r * UC + KM + lo] : 0.f;
                    x = (long long)rint((double)u * sc[j]);
                } else {
                    const float2 u = (r < R) ? ((const float2 *)U)[r * UC + KM + lo] : make_float2(0.f, 0.f);
                    const long long xr = (long long)rint((double)u.x * sc[j]), xi = (long long)rint((double)u.y * sc[j]);
                    x = CM == 1 ? xr : (CM == 2 ? xi : xr + xi);
                }
                Xc[r * K + lo] = (float)x;
            }
        }
        for (int ct = 0; ct < NQT; ct++) {
            uint32_t pl[QU] = {0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const long long r = rt * 16 + 4 * g + j;
                long long x;
                if constexpr (CM == 0) {
                    const float u = (r < R) ? U[r * UC + 16 * ct + lo] : 0.f;
                    x = (long long)rint((double)u * sc[j]);
                } else {
                    const float2 u = (r < R) ? ((const float2 *)U)[r * UC + 16 * ct + lo] : make_float2(0.f, 0.f);
                    const long long xr = (long long)rint((double)u.x * sc[j]), xi = (long long)rint((double)u.y * sc[j]);
                    x = CM == 1 ? xr : (CM == 2 ? xi : xr + xi);
                }
#pragma unroll
                for (int t = 0; t < QU; t++) {
                    const long long d = (t < QU - 1) ? bal8l(x) : x;
                    pl[t] |= ((uint32_t)d & 255u) << (8 * j);
                    x = (x - d) >> 8;
                }
            }
            uint32_t *o = Uq + ((rt * NQT + ct) * 64 + lane) * QU;
#pragma unroll
            for (int t = 0; t < QU; t++) o[t] = pl[t];
            if (Uf) {
                // row-major copy for the GEMM forward pass, transposed through this
                // wavefront's 768 bytes of LDS: chunk 16 ct + row of plane t = the 16
                // digits of row 16 rt + row at columns 16 ct .. 16 ct + 15
                uint8_t *tl = tile_t[wv];
#pragma unroll
                for (int t = 0; t < QU; t++)
#pragma unroll
                    for (int j = 0; j < 4; j++) tl[t * 256 + (4 * g + j) * 16 + lo] = (uint8_t)(pl[t] >> (8 * j));
                __builtin_amdgcn_wave_barrier();
                if (lane < 16 * QU) {
                    const int t = lane >> 4, row = lane & 15;
                    Uf[(rt * QU + t) * NQT * 16 + 16 * ct + row] = *(const uint4 *)(tl + t * 256 + row * 16);
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

// v -> forward B operands.  Thread per column col < UC: B[i][col] = v_w[k][i][j]
// (col = k M + j < K M), v_a[k][i] (col = K M + k), else 0; its scale and six digits
// per spin i (u = 0 least significant).  vfrag[(ct QV + u) 64 + lane] holds BPL bytes: spin k(l, e) = BPL (l/16) + e,
// column 16 ct + l%16.  colscale[col] = 2^(g - 46); bias[col] = v_b[k][j] or 0.
__global__ void k_q8_prep_v(SvfArgs a, int nsp, uint8_t *vfrag, double *colscale, double *bias)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int UC = a.Ucols;
    if (col >= UC) return;
    const int N = a.N, M = a.M, K = a.K;
    const long long KM = (long long)K * M;
    const double *v = (const double *)a.v;
    auto Bval = [&](int i) -> double {
        if (col < KM) return v[(long long)(col / M) * a.L + N + M + (long long)i * M + col % M];
        if (col < KM + K) return v[(long long)(col - KM) * a.L + i];
        return 0.0;
    };
    double mx = 0.0;
    for (int i = 0; i < N; i++) mx = fmax(mx, fabs(Bval(i)));
    int g = 0;
    if (mx > 0.0) (void)frexp(mx, &g);
    const double sc = ldexp(1.0, VBITS - g);
    colscale[col] = mx > 0.0 ? ldexp(1.0, g - VBITS) : 0.0;
    bias[col] = col < KM ? v[(long long)(col / M) * a.L + N + col % M] : 0.0;
    const int bpl = nsp / 4;
    const int ct = col / 16, lo = col % 16;
    for (int i = 0; i < nsp; i++) {
        long long q = (i < N) ? (long long)rint(Bval(i) * sc) : 0;
        int dg[QV];  // least significant first
#pragma unroll
        for (int u = 0; u < QV; u++) {
            const long long d = (u < QV - 1) ? bal8l(q) : q;
            dg[u] = (int)d;
            q = (q - d) >> 8;
        }
        const int lane = (i / bpl) * 16 + lo, e = i % bpl;
#pragma unroll
        for (int u = 0; u < QV; u++) vfrag[(((long long)ct * QV + u) * 64 + lane) * bpl + e] = (uint8_t)(dg[u] & 255);
    }
}

template <int NSP> struct FwdOp;
template <> struct FwdOp<32> {
    typedef long long T;
    static __device__ inline v4i mfma(T a, T b) { return __builtin_amdgcn_mfma_i32_16x16x32_i8(a, b, v4i{0, 0, 0, 0}, 0, 0, 0); }
};
template <> struct FwdOp<64> {
    typedef v4i T;
    static __device__ inline v4i mfma(T a, T b) { return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, v4i{0, 0, 0, 0}, 0, 0, 0); }
};

// +-1 bytes (0 for spins >= N or rows >= R) of spins [first, first + 4) of a row
__device__ inline uint32_t spin_bytes4(uint32_t bits, uint32_t valid)
{
    const uint32_t one = expand4(valid), neg = expand4(bits & valid);
    return one ^ (neg * 0xFEu);
}

// Forward pass.  Block (b, g) handles column tiles [g cgt, ...) (their v operands in
// LDS) and loops with its 4 or 8 wavefronts over 16-row tiles.  Lane l: A operand = spins
// BPL (l/16) .. + BPL of row 16 rt + l%16; result rows 16 rt + 4 (l/16) + e at column
// l%16 of each tile, matching the lane's elements of Uq.  Output wrow[g R + r].
template <int NSP>
__global__ __launch_bounds__(512) void k_q8_fwd(SvfArgs a, int cgt)
{
    using Op = FwdOp<NSP>;
    using AT = typename Op::T;
    constexpr int BPL = NSP / 4;
    extern __shared__ __align__(16) unsigned char smem_q[];
    const int N = a.N, K = a.K, NW = a.NW, UC = a.Ucols;
    const int NCT = UC / 16;
    const int ct0 = blockIdx.y * cgt;
    const int nct = (NCT - ct0 < cgt) ? NCT - ct0 : cgt;
    uint8_t *fr = smem_q;
    double *cs = (double *)(smem_q + (size_t)cgt * QV * 64 * BPL);
    double *bs = cs + 16 * cgt;
    {
        const uint4 *src = (const uint4 *)((const uint8_t *)a.vq + (size_t)ct0 * QV * 64 * BPL);
        uint4 *dst = (uint4 *)fr;
        const int n16 = nct * QV * 64 * BPL / 16;
        for (int q = threadIdx.x; q < n16; q += blockDim.x) dst[q] = src[q];
        const double *gcs = (const double *)((const uint8_t *)a.vq + (size_t)NCT * QV * 64 * 16);
        const double *gbs = gcs + UC;
        for (int q = threadIdx.x; q < 16 * nct; q += blockDim.x) {
            cs[q] = gcs[16 * ct0 + q];
            bs[q] = gbs[16 * ct0 + q];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, lo = lane & 15;
    const long long R = a.S * K;
    const long long ntile = (R + 15) / 16;
    const uint32_t *Uq = (const uint32_t *)a.Uq;
    double *wrow = (double *)a.wrow + (long long)blockIdx.y * R;
    const int wpb = blockDim.x >> 6;
    for (long long rt = (long long)blockIdx.x * wpb + wv; rt < ntile; rt += (long long)gridDim.x * wpb) {
        const long long ra = rt * 16 + lo;  // A operand row
        uint32_t w0 = 0, w1 = 0;
        if (ra < R) {
            w0 = a.spins[ra * NW];
            if (NW > 1) w1 = a.spins[ra * NW + 1];
        }
        AT av;
        {
            const int i0 = BPL * g;  // first spin of this lane
            uint32_t bytes[BPL / 4];
#pragma unroll
            for (int h = 0; h < BPL / 4; h++) {
                const int i = i0 + 4 * h;
                const uint32_t word = i < 32 ? w0 : w1;
                const uint32_t bits = (word >> (i & 31)) & 15u;
                const int nv = N - i;
                const uint32_t valid = (ra < R && nv > 0) ? (nv >= 4 ? 15u : ((1u << nv) - 1u)) : 0u;
                bytes[h] = spin_bytes4(bits, valid);
            }
            __builtin_memcpy(&av, bytes, sizeof(av));
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const uint3 *wrt = (const uint3 *)Uq + (rt * NCT + ct0) * 64 + lane;
        uint3 wn = wrt[0];  // digit planes of column tile c + 1 load while tile c is multiplied
        for (int c = 0; c < nct; c++) {
            const uint3 wq = wn;
            if (c + 1 < nct) wn = wrt[(c + 1) * 64];
            v4i y[QV];
#pragma unroll
            for (int u = 0; u < QV; u++) {
                const AT bv = *(const AT *)(fr + ((size_t)(c * QV + u) * 64 + lane) * BPL);
                y[u] = Op::mfma(av, bv);
            }
            const double csc = cs[16 * c + lo], bias = bs[16 * c + lo];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int h = (y[5][e] * 256 + y[4][e]) * 256 + y[3][e];
                const int l = (y[2][e] * 256 + y[1][e]) * 256 + y[0][e];
                const double yv = fma((double)h, 16777216.0, (double)l);
                const double z = fma(yv, csc, bias);
                acc[e] = fma((double)decode_x(wq.x, wq.y, wq.z, e), z, acc[e]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const double s = xor_sum16(acc[e]);
            const long long r = rt * 16 + 4 * g + e;
            if (lo == 0 && r < R) wrow[r] = s * a.escale[r];
        }
    }
}

// Chunks of the backward pass: 64 rows, row t = 16 q + 4 g + j of a chunk sits in
// byte j of dword q of lane group g (the k(l, e) of the 16x16x64 operands).  Block b
// of the backward grid takes chunks [b n64 / nblk, (b+1) n64 / nblk).
__device__ inline void block_chunks(long long n64, long long b, long long nb, long long &c0, long long &c1)
{
    c0 = n64 * b / nb;
    c1 = n64 * (b + 1) / nb;
}

// Spin masks, once per sample set: msk[(chunk NI + i) 4 + g] = uint4 over q of dwords
// with byte j = 0xff where sigma_{16 q + 4 g + j}[i] = -1 (0 for spins >= N, rows >= R).
// Thread per (chunk, i, g).
__global__ void k_q8_masks(const uint32_t *spins, long long R, int N, int NW, int NI, long long n64, uint4 *msk)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n64 * NI * 4) return;
    const int g = (int)(idx & 3), i = (int)((idx >> 2) % NI);
    const long long ch = idx / (4LL * NI);
    uint32_t d[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t bits = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const long long r = ch * 64 + 16 * q + 4 * g + j;
            if (r < R && i < N) bits |= ((spins[r * NW + (i >> 5)] >> (i & 31)) & 1u) << j;
        }
        d[q] = expand4(bits) * 0xffu;
    }
    msk[idx] = uint4{d[0], d[1], d[2], d[3]};
}

// Row weights of the backward pass, per call: omega_r = wt(s) 2^(e_r - 30) with
// wt = w_s, 1 or E_s (WSEL 0, 1, 2).  Block b (one per backward block): G_b from
// max |omega| over its rows (gblk[b]), then per chunk the digit dwords
// pd[(chunk QW + u) 4 + g] = uint4 over q of Omega = rint(omega 2^(46 - G_b)) (digit u,
// least significant first); and part[b][P] = sum of wt(s) over its rows with r % K == 0.
// With the c columns apart (a.qcols < a.Ucols; N <= 32, K <= 4) also the a block of
// part[b], sum_r sigma_r[i] X_c[r][k] Omega_r 2^(G - 46), on int8 MFMA as the backward
// pass forms the other blocks, T - 2 sum_r b_r[i] X_c Omega: per chunk the B operand
// holds the three digits d of X_c[r][k] as columns 3 k + d (built in the wavefront's
// LDS), the A operands are the weight digits u (masked by the spins, and unmasked for
// T), one int32 accumulator per (spin tile, u): QW (1 + NIT) MFMAs per chunk, every
// digit pair kept.  At the end each lane forms its column's FP64 sum over u (weight
// 256^(d + u)), the three digit columns of k are added, and the four wavefronts' sums
// are added in order.  CS = false: no a block (the kernel keeps its low register count);
// CS = true compiled for three wavefronts per SIMD (168 registers; unconstrained, 188 and two).
template <int WSEL, bool CS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CS ? 3 : 1))) void k_q8_omega(SvfArgs a, int *gblk, uint4 *pd, long long wss, long long q8ss,
                                                  long long pss, long long xss)
{
    // blockIdx.z: the digit set (weights a.w + z wss, block scales and digits + z q8ss
    // bytes, partial rows a.part + z pss doubles, c columns a.xc + z xss floats) when
    // several sets run in one launch
    a.w = (double *)a.w + blockIdx.z * wss;
    a.part = (double *)a.part + blockIdx.z * pss;
    gblk = (int *)((uint8_t *)gblk + blockIdx.z * q8ss);
    pd = (uint4 *)((uint8_t *)pd + blockIdx.z * q8ss);
    const float *xc = CS ? a.xc + blockIdx.z * xss : nullptr;
    __shared__ uint32_t stg[4][QW][4][4];  // per wavefront: a chunk's digits in the A layout
    __shared__ double red[4], rsum[4];
    __shared__ __align__(16) uint8_t xb[4][16][64];  // per wavefront: X_c digits, column 3 k + d, row t
    __shared__ __align__(16) uint32_t xs[4][64];     // and the chunk's spin words
    __shared__ double ared[4][128];                   // the wavefronts' a-block sums [k 32 + i]
    const int K = a.K;
    const long long R = a.S * K;
    const long long n64 = (R + 63) / 64;
    long long c0, c1;
    block_chunks(n64, blockIdx.x, gridDim.x, c0, c1);
    const double *w = (const double *)a.w;
    const double *es = a.escale;
    // sample of row r (R < 2^31, K <= 4): a shift for K = 1, 2, 4, a constant division for 3
    auto sample = [&](long long r) -> unsigned { return K == 3 ? (unsigned)r / 3u : (unsigned)r >> (K >> 1); };
    auto wt = [&](long long r) -> double {
        if constexpr (WSEL == 0) return w[sample(r)];
        else if constexpr (WSEL == 1) return 1.0;
        else return ((const nq64::cd *)a.E)[sample(r)].re;
    };
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double mx = 0.0, sw = 0.0;
    for (long long r = c0 * 64 + threadIdx.x; r < c1 * 64 && r < R; r += blockDim.x) {
        const double x = wt(r);
        mx = fmax(mx, fabs(x * es[r]));
        if (sample(r) * (unsigned)K == (unsigned)r) sw += x;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        mx = fmax(mx, __shfl_xor(mx, off));
        sw += __shfl_xor(sw, off);
    }
    if (lane == 0) {
        red[wv] = mx;
        rsum[wv] = sw;
    }
    __syncthreads();
    double m = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) m = fmax(m, red[q]);
    int G = 0;
    if (m > 0.0) (void)frexp(m, &G);  // m < 2^G
    if (threadIdx.x == 0) {
        gblk[blockIdx.x] = G;
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += rsum[q];
        ((double *)a.part)[(long long)blockIdx.x * (a.P + 1) + a.P] = t;
    }
    const double wsc = ldexp(1.0, WBITS - G);
    // wavefront wv takes chunks c0 + wv, c0 + wv + 4, ...: lane t computes row t's
    // digits into the wavefront's LDS slice, 24 lanes store them in the A layout (LDS
    // operations of one wavefront run in order; no block barrier)
    uint32_t(*st)[4][4] = stg[wv];
    const int N = a.N, nit = (N + 15) / 16;
    const int g4 = lane >> 4, lo = lane & 15;
    v4i cacc[2][QW], ctacc[QW];  // a block: masked (spin tiles) and unmasked (T) products per digit u
#pragma unroll
    for (int u = 0; u < QW; u++) {
        ctacc[u] = v4i{0, 0, 0, 0};
        cacc[0][u] = v4i{0, 0, 0, 0};
        cacc[1][u] = v4i{0, 0, 0, 0};
    }
    for (long long ch = c0 + wv; ch < c1; ch += 4) {
        // the chunk's global loads first, issued together: the row's weight and scale, its
        // K c columns (|X| < 2^24; rows past R hold zeros), its spin word
        const long long r = ch * 64 + lane;
        const double om = r < R ? wt(r) * es[r] : 0.0;
        float xv[4] = {0.f, 0.f, 0.f, 0.f};
        uint32_t sw0 = 0u;
        if constexpr (CS) {
#pragma unroll
            for (int k = 0; k < 4; k++) xv[k] = xc[r * K + (k < K ? k : 0)];
            sw0 = r < R ? a.spins[r * a.NW] : 0u;
        }
        {
            const int t = lane;
            if constexpr (CS) {  // digits of the c columns, the spin word
                xs[wv][t] = sw0;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    if (k < K) {
                        int x = (int)xv[k];
#pragma unroll
                        for (int d = 0; d < QU; d++) {
                            const int dg = (d < QU - 1) ? (int)(int8_t)(uint8_t)(x & 255) : x;
                            xb[wv][3 * k + d][t] = (uint8_t)(dg & 255);
                            x = (x - dg) >> 8;
                        }
                    }
                }
            }
            long long q = (long long)rint(om * wsc);
            const int qq = t >> 4, gg = (t >> 2) & 3, j = t & 3;
#pragma unroll
            for (int u = 0; u < QW; u++) {
                const long long d = (u < QW - 1) ? bal8l(q) : q;
                ((uint8_t *)&st[u][gg][qq])[j] = (uint8_t)(d & 255);
                q = (q - d) >> 8;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < QW * 4) {
            const int u = lane >> 2, g = lane & 3;
            pd[(ch * QW + u) * 4 + g] = uint4{st[u][g][0], st[u][g][1], st[u][g][2], st[u][g][3]};
        }
        if constexpr (CS) {
            // B: lane (g, col) dword q byte j = digit of row 16 q + 4 g + j at column col;
            // A: the digit-u dwords of the same rows (lanes of row group g), masked by the
            // spins 16 it + lo (byte j = 0xff where that spin of row 16 q + 4 g + j is down; spins
            // >= N give products that are not written)
            v4i bx;
            uint4 nm[2];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                bx[q] = lo < 3 * K ? *(const int *)&xb[wv][lo][16 * q + 4 * g4] : 0;
                const uint4 w4 = *(const uint4 *)&xs[wv][16 * q + 4 * g4];
#pragma unroll
                for (int it = 0; it < 2; it++) {
                    const int i = 16 * it + lo;
                    const uint32_t m = (((w4.x >> i) & 1u) | (((w4.y >> i) & 1u) << 8) | (((w4.z >> i) & 1u) << 16) |
                                        (((w4.w >> i) & 1u) << 24)) * 0xffu;
                    if (q == 0) nm[it].x = m;
                    else if (q == 1) nm[it].y = m;
                    else if (q == 2) nm[it].z = m;
                    else nm[it].w = m;
                }
            }
#pragma unroll
            for (int u = 0; u < QW; u++) {
                const uint32_t *pu = st[u][g4];
                const v4i aall = v4i{(int)pu[0], (int)pu[1], (int)pu[2], (int)pu[3]};
                ctacc[u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aall, bx, ctacc[u], 0, 0, 0);
#pragma unroll
                for (int it = 0; it < 2; it++) {
                    if (it < nit) {
                        const v4i av = v4i{(int)(pu[0] & nm[it].x), (int)(pu[1] & nm[it].y), (int)(pu[2] & nm[it].z),
                                           (int)(pu[3] & nm[it].w)};
                        cacc[it][u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bx, cacc[it][u], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if constexpr (CS) {
        // lane (g, col = 3 k + d), register e of spin tile it: spin 16 it + 4 g + e
        const double osc = ldexp(1.0, G - WBITS);
        const int d = lo % 3, k = lo / 3;
#pragma unroll
        for (int it = 0; it < 2; it++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                double h = 0.0;
#pragma unroll
                for (int u = QW - 1; u >= 0; u--) h = fma(h, 256.0, fma(-2.0, (double)cacc[it][u][e], (double)ctacc[u][0]));
                h = ldexp(h, 8 * d);
                const double h1 = __shfl(h, (lane & ~15) | ((lo + 1) & 15)), h2 = __shfl(h, (lane & ~15) | ((lo + 2) & 15));
                const int i = 16 * it + 4 * g4 + e;
                if (d == 0 && k < K && i < 32) ared[wv][k * 32 + i] = ((h + h1) + h2) * osc;
            }
        __syncthreads();
        for (int p = threadIdx.x; p < K * N; p += blockDim.x) {
            const int kk = p / N, i = p % N;
            double sum = 0.0;
            for (int q = 0; q < (int)(blockDim.x >> 6); q++) sum += ared[q][kk * 32 + i];
            ((double *)a.part)[(long long)blockIdx.x * (a.P + 1) + (long long)kk * a.L + i] = sum;
        }
    }
}

// Backward pass.  Wavefront wv of block (b, y, z) takes column tile ct = (4 / zw) y +
// wv / zw and spin tiles [z' NIT, z' NIT + NIT), z' = z zw + wv % zw, over the chunks of
// block row b (every slice forms T, slice 0 writes the b block); the zw slices of one
// column tile run in the same block, so the second reads the factor words from cache.
// Wavefronts share nothing (no LDS, no barriers).  Per chunk: B operand of factor digit t = dword t of
// the lane's digit planes of row tile q (q = 0..3); A operand of spin
// tile it, weight digit u: dword q = digits & mask.  Accumulator s = t + u collects
// sum_r b X Omega products of weight 256^s, tacc[s] the column sums T.  Each chunk's
// words, digits and masks load at the top of its iteration; three wavefronts per SIMD
// (166 registers, accumulators in VGPRs) cover each other's loads (measured faster than
// two wavefronts with the next chunk's operands prefetched into registers).  Writes
// part[b][0..P) for its columns: a and w blocks T - 2 sum b X Omega, b block T.
// blockIdx.z = set nzs + slice: digit set `set` (a.Uq + set uqss bytes, block scales and
// weight digits + set q8ss bytes, partial rows a.part + set pss doubles) when several
// sets run in one launch; `slice` takes the spin-tile slices of grids with nzs > 1.
template <int NIT>
__global__ __launch_bounds__(256, 3) void k_q8_bwd(SvfArgs a, const int *gblk, const uint4 *pd, const uint4 *msk,
                                                int zw, int nyl, int nzs, long long uqss, long long q8ss,
                                                long long pss)
{
    {
        const int set = blockIdx.z / nzs;
        a.Uq = (const uint8_t *)a.Uq + set * uqss;
        a.part = (double *)a.part + set * pss;
        gblk = (const int *)((const uint8_t *)gblk + set * q8ss);
        pd = (const uint4 *)((const uint8_t *)pd + set * q8ss);
    }
    const int N = a.N, M = a.M, K = a.K;
    const int NCT = a.qcols / 16;  // digit column tiles (the c columns apart: their a block comes from k_q8_omega)
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, lo = lane & 15;
    // 1-D grid of nb ny blocks: the ny column-tile blocks of one row range are adjacent
    // and the row ranges run from the last to the first, the reverse of the forward
    // pass, so the rows it read last (still in the device's last-level cache) come first
    const int ny = nyl;
    const long long lin = blockIdx.x;
    const long long nb = (long long)gridDim.x / ny;
    const long long bx = nb - 1 - lin / ny;
    const int by = (int)(lin % ny);
    const int ct = by * (4 / zw) + wv / zw;
    if (ct >= NCT) return;
    const int NIA = 16 * ((N + 15) / 16);  // spins of the mask array
    const int zs = (blockIdx.z % nzs) * zw + wv % zw;
    const int i0 = zs * NIT * 16;
    const long long R = a.S * K;
    const long long KM = (long long)K * M;
    const long long n64 = (R + 63) / 64;
    long long c0, c1;
    block_chunks(n64, bx, nb, c0, c1);
    const int G = gblk[bx];

    v4i acc[NIT][NS];
#pragma unroll
    for (int it = 0; it < NIT; it++)
#pragma unroll
        for (int s = 0; s < NS; s++) acc[it][s] = v4i{0, 0, 0, 0};
    v4i tacc[NS];  // T: the same products with the weight digits unmasked (all 16 rows equal)
#pragma unroll
    for (int s = 0; s < NS; s++) tacc[s] = v4i{0, 0, 0, 0};
    // Each wavefront's chunk operands go from global memory straight into its own LDS
    // slice (global_load_lds_dwordx4: lane l's 16 bytes at the instruction's LDS base + 16 l),
    // one chunk ahead, double-buffered: the four row tiles' digit words of column tile ct
    // (4 x 768 B, each contiguous in Uq), the weight digits (QW x 64 B) and the spin masks
    // of spins i0 .. i0 + 16 NIT (NIT x 1 KB).  No block barrier: a wavefront reads only
    // what it copied.  The two buffers are distinct arrays and the loop body is
    // instantiated per buffer, so that the compiler does not wait for the copies into one
    // before reading the other.
    constexpr int UQ_DW = 4 * 192, PD_DW = QW * 16, MK_DW = NIT * 256;
    constexpr int PW = UQ_DW + PD_DW + MK_DW;  // dwords per wavefront and buffer
    __shared__ __align__(16) uint32_t bqa[4 * PW], bqb[4 * PW];
    const uint4 *Uq4 = (const uint4 *)a.Uq;
    auto dma = [&](long long ch, uint32_t *b) {
        int ln = lane;  // opaque: per-lane addresses formed here, not held through the loop
        asm volatile("" : "+v"(ln));
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (ln < 48)
                __builtin_amdgcn_global_load_lds((const void *)(Uq4 + ((ch * 4 + q) * NCT + ct) * 48 + ln),
                                                 (__attribute__((address_space(3))) void *)(b + q * 192), 16, 0, 0);
        if (ln < QW * 4)
            __builtin_amdgcn_global_load_lds((const void *)(pd + ch * QW * 4 + ln),
                                             (__attribute__((address_space(3))) void *)(b + UQ_DW), 16, 0, 0);
#pragma unroll
        for (int it = 0; it < NIT; it++)
            __builtin_amdgcn_global_load_lds((const void *)(msk + (ch * NIA + i0 + 16 * it) * 4 + ln),
                                             (__attribute__((address_space(3))) void *)(b + UQ_DW + PD_DW + it * 256),
                                             16, 0, 0);
    };
    uint3 nw[4];
    uint4 nd[QW], nm[NIT];
    // chunks from the last to the first (the forward pass went the other way)
    auto body = [&](long long ch, const uint32_t *cur, uint32_t *nxt) {
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this chunk's copies are in LDS
#pragma unroll
        for (int q = 0; q < 4; q++) nw[q] = ((const uint3 *)(cur + q * 192))[lane];
#pragma unroll
        for (int u = 0; u < QW; u++) nd[u] = ((const uint4 *)(cur + UQ_DW))[u * 4 + g];
#pragma unroll
        for (int it = 0; it < NIT; it++) nm[it] = ((const uint4 *)(cur + UQ_DW + PD_DW))[(16 * it + lo) * 4 + g];
        // the next chunk's copies after this chunk's reads (a copy in flight before them
        // would make the compiler wait for it)
        if (ch - 1 >= c0) dma(ch - 1, nxt);
        const uint3 *cw = nw;
        const uint4 *cd = nd, *cm = nm;
        v4i bp[QU];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            bp[0][q] = (int)cw[q].x;
            bp[1][q] = (int)cw[q].y;
            bp[2][q] = (int)cw[q].z;
        }
#pragma unroll
        for (int u = 0; u < QW; u++) {
            const uint32_t pw[4] = {cd[u].x, cd[u].y, cd[u].z, cd[u].w};
            {
                const v4i aall = v4i{(int)pw[0], (int)pw[1], (int)pw[2], (int)pw[3]};
#pragma unroll
                for (int t = 0; t < QU; t++)
                    if (t + u >= QSKIP)
                        tacc[t + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aall, bp[t], tacc[t + u], 0, 0, 0);
            }
#pragma unroll
            for (int it = 0; it < NIT; it++) {
                const v4i av = v4i{(int)(pw[0] & cm[it].x), (int)(pw[1] & cm[it].y), (int)(pw[2] & cm[it].z),
                                   (int)(pw[3] & cm[it].w)};
#pragma unroll
                for (int t = 0; t < QU; t++)
                    if (t + u >= QSKIP)
                        acc[it][t + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bp[t], acc[it][t + u], 0, 0, 0);
            }
        }
    };
    uint32_t *const sa = bqa + wv * PW, *const sb = bqb + wv * PW;
    if (c1 - 1 >= c0) dma(c1 - 1, sa);
    for (long long ch = c1 - 1; ch >= c0; ch -= 2) {
        body(ch, sa, sb);
        if (ch - 1 >= c0) body(ch - 1, sb, sa);
    }
    double *part = (double *)a.part + bx * (a.P + 1);
    const double osc = ldexp(1.0, G - WBITS);
    const int col = 16 * ct + lo;
    int ts[NS];  // column sums: every row of the T products holds them; row 4 g of lane l
#pragma unroll
    for (int s = 0; s < NS; s++) ts[s] = tacc[s][0];
#pragma unroll
    for (int it = 0; it < NIT; it++) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            double hsum = 0.0;
#pragma unroll
            for (int s = NS - 1; s >= 0; s--)
                hsum = fma(hsum, 256.0, fma(-2.0, (double)acc[it][s][e], (double)ts[s]));
            const int i = i0 + 16 * it + 4 * g + e;
            if (i < N) {
                const double val = hsum * osc;
                if (col < KM) part[(col / M) * a.L + N + M + (long long)i * M + col % M] = val;
                else if (col < KM + K) part[(col - KM) * a.L + i] = val;
            }
        }
    }
    double hsum = 0.0;
#pragma unroll
    for (int s = NS - 1; s >= 0; s--) hsum = fma(hsum, 256.0, (double)ts[s]);
    if (g == 0 && col < KM && zs == 0) part[(col / M) * a.L + N + col % M] = hsum * osc;
}

// Complex factor rows: Umax[r] = max over columns of |re| + |im| (the scale of the three
// digit sets of row r).  16 lanes per row.
__global__ void k_q8c_rowmax(const float2 *U, long long R, int UC, float *Umax)
{
    const long long r = (long long)blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4);
    const int lo = threadIdx.x & 15;
    if (r >= R) return;
    float mx = 0.f;
    for (int c = lo; c < UC; c += 16) {
        const float2 u = U[r * UC + c];
        mx = fmaxf(mx, fabsf(u.x) + fabsf(u.y));
    }
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    if (lo == 0) Umax[r] = mx;
}

// Gauss's form of w_s = sum_kk U (Y + bias) for complex U and v, from the three forward
// runs' sample sums k1 = A B(Re v), k2 = R B(Im v - Re v), k3 = I B(Re v + Im v) (A = R + I):
// Re w_s = k1 - k3 - Re o^T v, Im w_s = k1 + k2 - Im o^T v; written in place as the
// backward runs' sample weights: set A takes Re w + Im w, set R Im w, set I Re w.
__global__ void k_q8c_combine3(double *w3, long long S, const double *ov)
{
    const double ovr = ov[0], ovi = ov[1];
    for (long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x; s < S; s += (long long)gridDim.x * blockDim.x) {
        const double k1 = w3[s], k2 = w3[S + s], k3 = w3[2 * S + s];
        const double wr = (k1 - k3) - ovr, wi = (k1 + k2) - ovi;
        w3[s] = wr + wi;
        w3[S + s] = wi;
        w3[2 * S + s] = wr;
    }
}

int nit_of(int N) { return (N + 15) / 16; }

template <int WSEL>
hipError_t launch_omega(const SvfArgs &a, const Q8Shape &sh, hipStream_t st, int nsets = 1, long long wss = 0,
                        long long q8ss = 0, long long pss = 0)
{
    int *gblk = (int *)a.q8w;
    uint4 *pd = (uint4 *)(gblk + sh.gblk_ints);
    const long long xss = (long long)(sh.xc_bytes / sizeof(float));
    if (a.qcols < a.Ucols)
        hipLaunchKernelGGL((k_q8_omega<WSEL, true>), dim3(sh.bwd_blocks, 1, nsets), dim3(256), 0, st, a, gblk, pd, wss,
                           q8ss, pss, xss);
    else
        hipLaunchKernelGGL((k_q8_omega<WSEL, false>), dim3(sh.bwd_blocks, 1, nsets), dim3(256), 0, st, a, gblk, pd,
                           wss, q8ss, pss, xss);
    return hipGetLastError();
}

// the backward products with the block scales and weight digits at a.q8w; nsets digit
// sets (strides uqss, q8ss bytes and pss doubles) in one launch
hipError_t launch_bwd_pass(const SvfArgs &a, const Q8Shape &sh, hipStream_t st, int nsets = 1, long long uqss = 0,
                           long long q8ss = 0, long long pss = 0)
{
    int *gblk = (int *)a.q8w;
    uint4 *pd = (uint4 *)(gblk + sh.gblk_ints);
    const uint4 *msk = (const uint4 *)a.q8m;
    const dim3 bb(256);
    const int nit = nit_of(a.N), nct = a.qcols / 16;
    const unsigned nb = (unsigned)sh.bwd_blocks, zn = (unsigned)nsets;
    if (nit == 1) {
        const int ny = (nct + 3) / 4;
        hipLaunchKernelGGL((k_q8_bwd<1>), dim3(nb * ny, 1, zn), bb, 0, st, a, gblk, pd, msk, 1, ny, 1, uqss, q8ss,
                           pss);
    } else if (nit == 2) {  // one slice of two spin tiles
        const int ny = (nct + 3) / 4;
        hipLaunchKernelGGL((k_q8_bwd<2>), dim3(nb * ny, 1, zn), bb, 0, st, a, gblk, pd, msk, 1, ny, 1, uqss, q8ss,
                           pss);
    } else if (nit == 4) {  // two slices of two spin tiles, in one block
        const int ny = (nct + 1) / 2;
        hipLaunchKernelGGL((k_q8_bwd<2>), dim3(nb * ny, 1, zn), bb, 0, st, a, gblk, pd, msk, 2, ny, 1, uqss, q8ss,
                           pss);
    } else {
        const int ny = (nct + 3) / 4;
        hipLaunchKernelGGL((k_q8_bwd<1>), dim3(nb * ny, 1, (unsigned)nit * zn), bb, 0, st, a, gblk, pd, msk, 1, ny,
                           nit, uqss, q8ss, pss);
    }
    return hipGetLastError();
}

template <int WSEL>
hipError_t launch_bwd(const SvfArgs &a, const Q8Shape &sh, hipStream_t st)
{
    const hipError_t e = launch_omega<WSEL>(a, sh, st);
    return e != hipSuccess ? e : launch_bwd_pass(a, sh, st);
}

} // namespace

// Applies for real weights and N <= 64.  Forward: column groups of cgt tiles whose v
// operands fit in 64 KB of LDS.  Backward: GS <= 8 wavefronts (one column tile each)
// per block, groups balanced; at most MAX_BLOCK_CHUNKS chunks per block (int32 sums
// stay exact).
// Work buffers: q8w (block scales, weight digits; per call), q8m (spin masks; per
// sample set).
Q8Shape nq_q8_shape(int N, int M, int K, long long R, bool cplx)
{
    Q8Shape sh{};
    if (N > 64 || K > 4 || R >= (1LL << 31)) return sh;  // k_q8_omega takes 32-bit row indices, K <= 4
    const int UC = 16 * ((K * M + K + 15) / 16);
    const int nct = UC / 16;
    sh.nsp = N <= 32 ? 32 : 64;
    const int bpl = sh.nsp / 4;
    const size_t per_tile = (size_t)QV * 64 * bpl + 2 * 16 * sizeof(double);
    int cgt = (int)(65536 / per_tile);
    if (cgt > nct) cgt = nct;
    sh.ucols = UC;
    sh.fwd_cgt = cgt;
    sh.fwd_cgs = (nct + cgt - 1) / cgt;
    sh.fwd_lds = (size_t)cgt * per_tile;
    sh.fwd_blocks = 1280;  // measured: 5 wavefronts per SIMD
    sh.fwd_threads = 256;
    if ((160 * 1024 / sh.fwd_lds) * 4 < 16) {  // LDS allows < 4 wavefronts per SIMD: 8 per block
        sh.fwd_threads = 512;
        sh.fwd_blocks = 640;
    }
    const int ngrp = (nct + 7) / 8;
    sh.bwd_cgs = ngrp;
    sh.bwd_waves = (nct + ngrp - 1) / ngrp;
    const long long n64 = (R + 63) / 64;
    // about 128 chunks per block, at least 256 blocks (measured: fewer, longer row ranges
    // write and reduce fewer partial rows; complex N=32 M=64 K=2: 1.19 ms per S*v with 1024
    // blocks, 1.16 with 512, 1.13 with 256; N=64 M=128 K=4 keeps 1024)
    long long nb = n64 / 128;
    if (nb < 256) nb = 256;
    if (nb > n64) nb = n64 > 0 ? n64 : 1;
    while (nb * MAX_BLOCK_CHUNKS < n64) nb *= 2;
    sh.bwd_blocks = (int)nb;
    const char *e = getenv("NQES_Q8_BWD_BLOCKS");
    if (e && *e && atoi(e) > 0 && (long long)atoi(e) * MAX_BLOCK_CHUNKS >= n64) sh.bwd_blocks = atoi(e);
    e = getenv("NQES_Q8_FWD_BLOCKS");
    if (e && *e && atoi(e) > 0) sh.fwd_blocks = atoi(e);
    sh.rpad = n64 * 64;
    sh.vq_bytes = (size_t)nct * QV * 64 * 16 + 2 * (size_t)UC * sizeof(double);
    // c columns apart (GEMM form, K in {1, 2, 4}): when the K c columns would open a column
    // tile of their own (K M a multiple of 16), the digits cover the K M tanh columns only
    // and the c columns stay integers (xc, rpad K floats per digit set), their products in
    // FP64 (forward: in k_q8_fwdg before the spin signs; backward: the a block in
    // k_q8_omega).  NQES_Q8_CSPLIT=0 keeps them as digit columns.
    const int QC = 16 * ((K * M + 15) / 16);
    const char *es = getenv("NQES_Q8_CSPLIT");
    const bool csplit = N <= 32 && (K == 1 || K == 2 || K == 4) && QC < UC && !(es && strcmp(es, "0") == 0);
    // GEMM forward pass: N <= 32, at most 4 groups of 64 digit columns (their v digits in
    // 52 KB of LDS); a row-major copy of the factor digits (3 R qcols bytes)
    sh.ncg64 = ((csplit ? QC : UC) + 63) / 64;
    sh.nst = (N + 15) / 16;
    const char *eg = getenv("NQES_Q8_FWD");  // "tiles": the tiled forward pass instead
    sh.gemm = N <= 32 && sh.ncg64 <= 4 && !(eg && strcmp(eg, "tiles") == 0);
    sh.qcols = sh.gemm && csplit ? QC : UC;
    sh.xc_bytes = sh.qcols < UC ? (size_t)n64 * 64 * K * sizeof(float) : 0;
    if (!sh.gemm) sh.ncg64 = (UC + 63) / 64;
    // v digits, their scale and (c columns apart) the K x 32 FP64 B values of the c columns
    sh.qg_bytes = (size_t)sh.ncg64 * (sh.nst * QV + 1) * 64 * 16 + 16 + (sh.xc_bytes ? (size_t)K * 32 * 8 : 0);
    sh.uf_bytes = sh.gemm ? (size_t)(n64 * 4) * QU * sh.qcols * 16 : 0;
    if (sh.gemm && sh.qg_bytes > sh.vq_bytes) sh.vq_bytes = sh.qg_bytes;
    if (sh.gemm) {
        sh.fwd_blocks = sh.nst * sh.ncg64 > 6 ? 512 : 768;  // 2 or 3 wavefronts per SIMD, as k_q8_fwdg's bounds
        e = getenv("NQES_Q8_FWD_BLOCKS");
        if (e && *e && atoi(e) > 0) sh.fwd_blocks = atoi(e);
    }
    sh.gblk_ints = (sh.bwd_blocks + 3) / 4 * 4;  // keeps pn 16-byte aligned
    sh.w_bytes = (size_t)sh.gblk_ints * 4 + (size_t)n64 * QW * 4 * 16;
    sh.m_bytes = (size_t)n64 * 16 * ((N + 15) / 16) * 4 * 16;
    // one-pass form: real weights, N <= 32, K | 4, K M <= 128 (the int32 accumulators of a
    // wavefront's column tiles); NQES_Q8_FUSED=0 keeps the two passes
    const int fz_ncg = (K * M + 63) / 64, fz_nst = (N + 15) / 16;
    const char *ef = getenv("NQES_Q8_FUSED");
    sh.fused = !cplx && N <= 32 && (K == 1 || K == 2 || K == 4) && fz_ncg <= 2 &&
               !(ef && strcmp(ef, "0") == 0);
    if (sh.fused) {
        sh.fz_ncg = fz_ncg;
        sh.fz_nst = fz_nst;
        // eight-wavefront blocks (even fz_ncg) take super-chunks of two 64-row chunks: rows
        // padded to a multiple of 128; one block per CU (LDS), else two
        const int nkc = fz_ncg % 2 == 0 ? 2 : 1;
        const long long n64f = (n64 + nkc - 1) / nkc * nkc, nsc = n64f / nkc;
        sh.rpad = n64f * 64;
        long long nbf = nkc == 2 ? 256 : 512;
        if (nbf > nsc) nbf = nsc > 0 ? nsc : 1;
        while (nbf * MAX_BLOCK_CHUNKS < n64f) nbf *= 2;
        e = getenv("NQES_Q8_FUSED_BLOCKS");
        if (e && *e && atoi(e) > 0 && atoi(e) <= nsc && (long long)atoi(e) * MAX_BLOCK_CHUNKS >= n64f) nbf = atoi(e);
        sh.fz_blocks = (int)nbf;
        if (sh.bwd_blocks < sh.fz_blocks) sh.bwd_blocks = sh.fz_blocks;  // partial rows allocated for the larger
        sh.gemm = 0;
        sh.uf_bytes = 0;
        sh.qcols = UC;
        sh.xc_bytes = 0;
        sh.ug_bytes = (size_t)sh.rpad * QU * 64 * fz_ncg;
        // forward B operands and the scale of v
        const size_t qf = (size_t)fz_ncg * (fz_nst * QV + 1) * 64 * 16 + 16;
        if (qf > sh.vq_bytes) sh.vq_bytes = qf;
        // the quantised c columns and a copy of the spin words padded to rpad rows in the
        // mask buffer
        const size_t xc = (size_t)sh.rpad * K * 4 + (size_t)sh.rpad * 4;
        if (xc > sh.m_bytes) sh.m_bytes = xc;
    }
    sh.ok = 1;
    return sh;
}

namespace {

// quantise a.U into a.Uq (and a.Uf for the GEMM form) with digit mode CM; masks too if asked
template <int CM>
hipError_t q8_pack_rows(const SvfArgs &a, const Q8Shape &sh, bool masks, hipStream_t st)
{
    const long long R = a.S * a.K;
    const long long ntile = sh.rpad / 16;
    long long blocks = (ntile + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL((k_q8_pack<CM>), dim3((unsigned)blocks), dim3(256), 0, st, (const float *)a.U, R, ntile,
                       a.Ucols, (uint32_t *)a.Uq, (double *)a.escale, sh.gemm ? (uint4 *)a.Uf : (uint4 *)nullptr,
                       a.Umax, sh.qcols, sh.qcols < a.Ucols ? a.xc : (float *)nullptr, a.K * a.M, a.K);
    if (masks) {
        const int NI = 16 * nit_of(a.N);
        const long long n64 = sh.rpad / 64, nthr = n64 * NI * 4;
        hipLaunchKernelGGL(k_q8_masks, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, st, a.spins, R, a.N,
                           a.NW, NI, n64, (uint4 *)a.q8m);
    }
    return hipGetLastError();
}

// forward pass: a.w[s] = sum over the sample's rows of X (sigma^T Q + bias) / f - *a.ov
hipError_t q8_forward(SvfArgs a, const Q8Shape &sh, hipStream_t st)
{
    const int UC = a.Ucols, nct = UC / 16;
    if (sh.gemm) {
        a.ncg = 1;
        const hipError_t e = nq_q8_fwd_gemm(a, sh, st);
        if (e != hipSuccess) return e;
    } else {
        a.ncg = sh.fwd_cgs;
        uint8_t *vfrag = (uint8_t *)a.vq;
        double *colscale = (double *)(vfrag + (size_t)nct * QV * 64 * 16);
        hipLaunchKernelGGL(k_q8_prep_v, dim3((UC + 63) / 64), dim3(64), 0, st, a, sh.nsp, vfrag, colscale,
                           colscale + UC);
        const dim3 grid(sh.fwd_blocks, sh.fwd_cgs);
        if (sh.nsp == 32) hipLaunchKernelGGL((k_q8_fwd<32>), grid, dim3(sh.fwd_threads), sh.fwd_lds, st, a, sh.fwd_cgt);
        else hipLaunchKernelGGL((k_q8_fwd<64>), grid, dim3(sh.fwd_threads), sh.fwd_lds, st, a, sh.fwd_cgt);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (!(sh.gemm && (a.K == 1 || a.K == 2 || a.K == 4)))  // else the forward pass wrote w
        e = nq_launch_svf_combine(false, a, st);
    return e;
}

// backward pass with sample weights a.w: out[p] = sum_r sigma X omega, out[P] = sum_s w_s
hipError_t q8_backward(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out, hipStream_t st)
{
    hipError_t e = launch_bwd<0>(a, sh, st);
    if (e != hipSuccess) return e;
    return nq_chunk_reduce(false, a.part, a.P + 1, a.P + 1, sh.bwd_blocks, scratch, out, st);
}

} // namespace

hipError_t nq_q8_pack(const SvfArgs &a, const Q8Shape &sh, hipStream_t st)
{
    if (sh.fused) {
        switch (a.K) {
        case 1: return nq_q8f_pack_k1(a, sh, st);
        case 2: return nq_q8f_pack_k2(a, sh, st);
        default: return nq_q8f_pack_k4(a, sh, st);
        }
    }
    return q8_pack_rows<0>(a, sh, true, st);
}

hipError_t nq_q8_apply(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out, hipStream_t st)
{
    if (sh.fused) {
        switch (a.K) {
        case 1: return nq_q8f_apply_k1(a, sh, scratch, out, st);
        case 2: return nq_q8f_apply_k2(a, sh, scratch, out, st);
        default: return nq_q8f_apply_k4(a, sh, scratch, out, st);
        }
    }
    hipError_t e = q8_forward(a, sh, st);
    return e != hipSuccess ? e : q8_backward(a, sh, scratch, out, st);
}

hipError_t nq_q8_moments(const SvfArgs &a0, const Q8Shape &sh, const SvfShape &svf, void *part, void *scratch,
                         void *tmp, void *so, void *ss, void *sf, hipStream_t st)
{
    SvfArgs a = a0;
    a.part = part;
    double *r1 = (double *)tmp, *rE = r1 + (a.P + 1);
    if (sh.fused) {
        hipError_t e = hipSuccess;
        for (int wsel = 1; wsel <= 2 && e == hipSuccess; wsel++) {
            void *o = wsel == 1 ? (void *)r1 : (void *)rE;
            if (a.K == 1) e = nq_q8f_moment_k1(a, sh, wsel, part, scratch, o, st);
            else if (a.K == 2) e = nq_q8f_moment_k2(a, sh, wsel, part, scratch, o, st);
            else e = nq_q8f_moment_k4(a, sh, wsel, part, scratch, o, st);
        }
        if (e != hipSuccess) return e;
        return nq_launch_svf_sq_pack(false, a0, svf, part, scratch, tmp, so, ss, sf, st);
    }
    hipError_t e = launch_bwd<1>(a, sh, st);
    if (e == hipSuccess) e = nq_chunk_reduce(false, part, a.P + 1, a.P + 1, sh.bwd_blocks, scratch, r1, st);
    if (e == hipSuccess) e = launch_bwd<2>(a, sh, st);
    if (e == hipSuccess) e = nq_chunk_reduce(false, part, a.P + 1, a.P + 1, sh.bwd_blocks, scratch, rE, st);
    if (e != hipSuccess) return e;
    return nq_launch_svf_sq_pack(false, a0, svf, part, scratch, tmp, so, ss, sf, st);
}

// Complex weights: row maxima of |re| + |im|, the digit sets X_A = X_R + X_I, X_R, X_I in
// slices 0, 1, 2 of Uq / Uf (one scale per row, shared), and the spin masks once.
hipError_t nq_q8c_pack(const SvfArgs &a0, const Q8Shape &sh, hipStream_t st)
{
    SvfArgs a = a0;
    const long long R = a.S * a.K;
    hipLaunchKernelGGL(k_q8c_rowmax, dim3((unsigned)((R + 15) / 16)), dim3(256), 0, st, (const float2 *)a.U, R,
                       a.Ucols, (float *)a.Umax);
    const size_t qs = (size_t)sh.rpad * a.Ucols * QU, fs = sh.uf_bytes;
    hipError_t e = hipGetLastError();
    for (int t = 0; t < 3 && e == hipSuccess; t++) {
        a.Uq = (const uint8_t *)a0.Uq + t * qs;
        a.Uf = sh.gemm ? (const void *)((const uint8_t *)a0.Uf + t * fs) : nullptr;
        a.xc = sh.xc_bytes ? a0.xc + t * (sh.xc_bytes / sizeof(float)) : nullptr;
        if (t == 0) e = q8_pack_rows<3>(a, sh, true, st);
        else if (t == 1) e = q8_pack_rows<1>(a, sh, false, st);
        else e = q8_pack_rows<2>(a, sh, false, st);
    }
    return e;
}

// S*v of complex weights on the int8 form, three real products per complex product
// (Gauss): forward runs A on Re v, R on Im v - Re v, I on Re v + Im v; k_q8c_combine3;
// backward runs A with Re w + Im w, R with Im w, I with Re w.  The digit sets satisfy
// X_A = X_R + X_I exactly, so the operator is the exact complex product of the quantised
// factors, up to the quantisation of v and w (2^-47 of their scale), as the real form.
// The per-set small kernels run once for all three sets: the v digits (GEMM form), the row
// weights and block scales, and the ordered reduction of the partial rows; the buffers
// a.vq, a.q8w and a.part hold three sets (nq_q8c_sizes).  Every set's sums are the same
// operations in the same order as when the sets run one after the other.
hipError_t nq_q8c_apply(const SvfArgs &a0, const Q8Shape &sh, const double *vin, double *w3, void *scratch,
                        double *out, hipStream_t st)
{
    SvfArgs a = a0;
    const long long P = a.P, S = a.S;
    const size_t qs = (size_t)sh.rpad * a.Ucols * QU, fs = sh.uf_bytes;
    const long long pss = (long long)sh.bwd_blocks * (P + 1);  // partial rows of one set (doubles)
    hipError_t e = hipSuccess;
    a.v = vin;
    a.ov = vin + 3 * P + 2;  // zero: o^T v enters in the combine
    if (sh.gemm) e = nq_q8_prep_vg_sets(a, sh, 3, P, sh.vq_bytes, st);
    if (sh.gemm && (a.K == 1 || a.K == 2 || a.K == 4)) {  // the GEMM passes write w: one launch
        a.Uf = a0.Uf;
        a.vq = a0.vq;
        a.w = w3;
        if (e == hipSuccess)
            e = nq_q8_fwd_gemm(a, sh, st, false, false, 3, (long long)fs, (long long)sh.vq_bytes, S,
                               (long long)(sh.xc_bytes / sizeof(float)));
    } else {
        for (int t = 0; t < 3 && e == hipSuccess; t++) {
            a.Uq = (const uint8_t *)a0.Uq + t * qs;
            a.Uf = sh.gemm ? (const void *)((const uint8_t *)a0.Uf + t * fs) : nullptr;
            a.v = vin + t * P;
            a.w = w3 + t * S;
            a.vq = (uint8_t *)a0.vq + t * sh.vq_bytes;
            if (sh.gemm) {
                e = nq_q8_fwd_gemm(a, sh, st, false, false);
                if (e == hipSuccess) e = nq_launch_svf_combine(false, a, st);
            } else {
                e = q8_forward(a, sh, st);
            }
        }
    }
    if (e != hipSuccess) return e;
    const long long nb = (S + 255) / 256 < 4096 ? (S + 255) / 256 : 4096;
    hipLaunchKernelGGL(k_q8c_combine3, dim3((unsigned)nb), dim3(256), 0, st, w3, S, vin + 3 * P);
    e = hipGetLastError();
    a = a0;
    a.w = w3;
    if (e == hipSuccess) e = launch_omega<0>(a, sh, st, 3, S, (long long)sh.w_bytes, pss);
    if (e == hipSuccess) e = launch_bwd_pass(a, sh, st, 3, (long long)qs, (long long)sh.w_bytes, pss);
    if (e == hipSuccess)
        e = nq_chunk_reduce_sets((const double *)a0.part, P + 1, P + 1, sh.bwd_blocks, (double *)scratch, out, 3, pss,
                                 P + 1, st);
    return e;
}

// device bytes the three-set complex S*v needs in a.vq, a.q8w, a.part and its scratch
void nq_q8c_sizes(const Q8Shape &sh, long long P, size_t *vq, size_t *q8w, size_t *part, size_t *scratch)
{
    *vq = 3 * sh.vq_bytes;
    *q8w = 3 * sh.w_bytes;
    *part = (size_t)3 * sh.bwd_blocks * (P + 1) * 8;
    *scratch = (size_t)3 * ((sh.bwd_blocks + 31) / 32) * (P + 1) * 8;
}
