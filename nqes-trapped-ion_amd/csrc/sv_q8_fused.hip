// sv_q8_fused.hip -- the int8 S*v of the mixed path (real weights) in one pass over the
// factor digits, for N <= 32 and K in {1, 2, 4}.  Reference: the matrix-free S*v of
// rbm.cc:661-684 (w_s = O_s^T v - o^T v, then sum_s conj(O_s) w_s) with O_s in the
// factored form sum_kk sigma^kk (x) c tanh(theta) (sv_factored.hip, sv_q8.hip).
//
// The two-pass int8 form reads the factor digits twice per call, a row-major copy in the
// forward GEMM (sv_q8_gemm.hip) and 16x16 tiles in the backward pass (sv_q8.hip).  Here
// a block takes super-chunks of 64-row chunks, copied from HBM into LDS (its digits as
// row-major 64 x 16-byte matrices, one per column tile, digit plane and chunk) while the
// previous super-chunk is processed.  Each of its wavefronts
//   1. runs the forward GEMM of one 16-row tile, P_ri = sum_c X_rc Q_ci, on
//      v_mfma_i32_16x16x64_i8 with its A operands from those matrices (layout Ug below),
//      and forms the rows' weights omega_r = (sum_kk w_row - o^T v) / f_r;
//   2. stores the rows' weight digits and spin masks in LDS;
// and after one block barrier the wavefronts run the chunks' backward GEMM
//   out[i][col] = sum_r sigma_r[i] X_r[col] Omega_r = T[col] - 2 sum_r b_r[i] X_r[col] Omega_r
// from LDS, each for its column tiles, into int32 accumulators.  The backward B operands
// (16 rows of one column per lane) come out of the row-major matrices through the LDS
// transpose load ds_read_b64_tr_b8, so no VALU work goes into the transpose.  The digits
// are read from HBM once per call.
//
// Columns.  The MFMAs cover the K M tanh columns of the factor rows (64-column groups).
// The K columns c_{s,k,kk} (the a block) are added in FP64 on the VALU from their
// quantised values Xc (the same integers as the digits would hold): in the forward pass
// Xc[r][k] times v_a[k][i] before the spin signs, in the backward pass
// sum_r sigma_r[i] Xc[r][k] omega_r per lane.  T[col] (the b block, and the spin-free
// half of the identity above) takes three MFMAs per column tile and chunk: the A operand
// holds weight digit u of the rows in its row u, so one product per factor digit t gives
// every digit pair (t, u).
//
// Weight scale.  Omega = rint(omega 2^(46 - G)) needs |omega| < 2^G.  The block keeps a
// running G: a chunk whose largest |omega| does not fit converts the accumulators to FP64
// into the block's partial row (at the old scale), clears them and raises G to that
// chunk's exponent + 2 (a few times per block).  Both passes leave out the digit pairs
// of weight 256^0 and 256^1 (QSKIP), as the two-pass form does.
//
// Layout Ug (written once per sample set by k_q8_packg): per super-chunk of 16 W rows
// (W wavefronts per block), the block's LDS image XT of its digits:
// [column tile][digit plane t][64-row chunk] row-major 64 x 16-byte matrices, row
// 16 q + lo of chunk kc = row lo of the super-chunk's row tile 4 kc + q.  The kernel copies
// a super-chunk into LDS as it is (global_load_lds_dwordx4) one super-chunk ahead.  The
// forward A operand of lane (g, lo) is row lo's 16 bytes of column tile 4 cg + g (columns
// 64 cg + 16 g + e, e = 0..15); the v digits (fz_v_operands) use the same column order.
//
// LDS per block (static arrays): the forward B operands (NCG (6 NST + 1) KB), and two
// buffers (super-chunk parity), each the digit image XT, the weight digits
// [chunk][g'][u][row tile] dwords, the spin masks [st][g'][chunk][i][row tile] dwords, the
// wavefronts' largest |omega|, and the copied row scales, c-column integers and spin words.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cplx64.hpp"
#include "kernels.hpp"
#include "sv_q8_common.hpp"

#ifndef NQES_FZ_K
#define NQES_FZ_K 2
#endif

namespace {

constexpr int FZ_RED = 128;  // doubles of the end-of-block reductions per wavefront (NST K 16)

// wavefronts per block: eight (128-row super-chunks, one column tile per wavefront in the
// backward products) when the column tiles split evenly, else four
__host__ __device__ constexpr int fz_waves(int ncg) { return (ncg % 2 == 0) ? 8 : 4; }

// four 0/1 bits -> four 0/1 bytes
__device__ inline uint32_t fz_expand4(uint32_t b) { return ((b & 15u) * 0x00204081u) & 0x01010101u; }

__device__ inline double fz_shfl(double x, int src)
{
    const long long b = __double_as_longlong(x);
    const int lo = __shfl((int)b, src), hi = __shfl((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL>
__device__ inline double fz_dpp(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// the four sums h[e] over the 16 lanes of a row at once (lane l returns the sum of h[l / 4])
__device__ inline double fz_sum16x4(const double (&h)[4], int l)
{
    const bool lowhalf = l < 8, lowquad = (l & 7) < 4;
    const double k0 = lowhalf ? h[0] : h[2], k1 = lowhalf ? h[1] : h[3];
    const double s0 = lowhalf ? h[2] : h[0], s1 = lowhalf ? h[3] : h[1];
    const double a0 = k0 + fz_dpp<0x140>(s0), a1 = k1 + fz_dpp<0x140>(s1);
    double b = (lowquad ? a0 : a1) + fz_dpp<0x141>(lowquad ? a1 : a0);
    b += fz_dpp<0x4E>(b);
    b += fz_dpp<0xB1>(b);
    return b;
}

typedef int v2i __attribute__((ext_vector_type(2)));

// ds_read_b64_tr_b8 at an LDS address (8-byte aligned)
__device__ inline v2i lds_tr8(const uint32_t *p)
{
    return __builtin_amdgcn_ds_read_tr8_b64_v2i32((__attribute__((address_space(3))) v2i *)(p));
}

__device__ inline v4i as_v4i(uint4 x)
{
    v4i r;
    __builtin_memcpy(&r, &x, 16);
    return r;
}

struct FzArgs {
    int N, M, K, L, KM;
    long long S, R, P, n64;
    const uint32_t *spins;  // one word per row (N <= 32)
    const uint32_t *spad;   // the same, rpad rows (zeros past R), written by the pack
    const uint4 *Ug;
    const float *Xc;        // [rpad][K] quantised c columns
    const double *escale;   // [rpad] 1 / f_r
    const double *v;        // v (the a blocks)
    const double *ov;       // o^T v
    const nq64::cd *E;      // local energies (wsel 2)
    double *part;           // gridDim.x x (P + 1)
    int wsel;               // MODE 1: 1 -> weights 1, 2 -> Re E_s
};

// ---- pack: factor rows U (FP32 [R][UC], or (re, im) pairs for CM > 0) -> Ug, Xc, escale.
// One wavefront per 16-row tile; the rows go through LDS (stride UC FW + 1 floats).
// Scale f_r = XMAX / 2^e_r (CM = 0) or (XMAX - 1) / 2^e_r (complex digit sets) with
// 2^e_r > Umax[r], as the two-pass pack.
template <int CM>
__global__ __launch_bounds__(256) void k_q8_packg(const float *U, long long R, long long ntile, int UC, int KM, int K,
                                                  int NCG, uint4 *Ug, float *Xc, double *escale, const float *Umax,
                                                  const uint32_t *spins, uint32_t *spad)
{
    const int W = fz_waves(NCG), NKC = W / 4;
    const long long XT_U4 = 4LL * NCG * QU * NKC * 64;  // uint4 of one super-chunk's digit image
    extern __shared__ float lds_pack[];
    constexpr int FW = CM ? 2 : 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, lo = lane & 15;
    const int rs = UC * FW + 1;  // LDS row stride
    float *tu = lds_pack + (size_t)wv * 16 * rs;
    for (long long rt = (long long)blockIdx.x * 4 + wv; rt < ntile; rt += (long long)gridDim.x * 4) {
        const long long r = rt * 16 + lo;
        const float mx = r < R ? Umax[r] : 0.f;
        if (g == 0) spad[r] = r < R ? spins[r] : 0u;
        int e = 0;
        if (mx > 0.f) (void)frexpf(mx, &e);  // mx < 2^e
        const double sc = ldexp(CM ? XMAX - 1.0 : XMAX, -e);
        if (g == 0) escale[r] = (r < R && mx > 0.f) ? 1.0 / sc : 0.0;
        __builtin_amdgcn_wave_barrier();
        for (int row = 0; row < 16; row++) {
            const long long rr = rt * 16 + row;
            const float *src = U + rr * (long long)UC * FW;
            for (int q = lane; q < UC * FW; q += 64) tu[row * rs + q] = rr < R ? src[q] : 0.f;
        }
        __builtin_amdgcn_wave_barrier();
        auto xval = [&](int col) -> long long {
            if constexpr (CM == 0) {
                return (long long)rint((double)tu[lo * rs + col] * sc);
            } else {
                const long long xr = (long long)rint((double)tu[lo * rs + 2 * col] * sc);
                const long long xi = (long long)rint((double)tu[lo * rs + 2 * col + 1] * sc);
                return CM == 1 ? xr : (CM == 2 ? xi : xr + xi);
            }
        };
        // lane (g, lo): row lo of the tile, the 16 columns of column tile 4 cg + g, into the
        // super-chunk's digit image at [column tile][t][chunk][row 16 q + lo]
        const long long sck = rt / W;  // the super-chunk
        const int wq = (int)(rt % W), kc = wq >> 2, q = wq & 3;
        for (int cg = 0; cg < NCG; cg++) {
            uint32_t pl[QU][4];
#pragma unroll
            for (int t = 0; t < QU; t++)
#pragma unroll
                for (int d = 0; d < 4; d++) pl[t][d] = 0u;
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    const int col = 64 * cg + 16 * g + 4 * b + d;
                    long long x = col < KM ? xval(col) : 0;
#pragma unroll
                    for (int t = 0; t < QU; t++) {
                        const long long dg = (t < QU - 1) ? bal8l(x) : x;
                        pl[t][b] |= ((uint32_t)dg & 255u) << (8 * d);
                        x = (x - dg) >> 8;
                    }
                }
            const int ct = 4 * cg + g;
#pragma unroll
            for (int t = 0; t < QU; t++)
                Ug[sck * XT_U4 + ((long long)(ct * QU + t) * NKC + kc) * 64 + 16 * q + lo] =
                    uint4{pl[t][0], pl[t][1], pl[t][2], pl[t][3]};
        }
        if (g == 0)
            for (int k = 0; k < K; k++) Xc[r * K + k] = (float)xval(KM + k);
    }
}

// ---- v -> forward B operands, formed in LDS by every block of the one-pass kernel (34 KB
// of v read from L2 per block; a separate kernel and its launch cost more).
// qs[(cg NSLOT + slot) 64 + lane]: lane (g, lo) byte e = digit u of v at column
// 64 cg + 16 g + e (the A operands' order: lane group g holds column tile 4 cg + g) for
// spin 16 st + lo (slot st QV + u), or digit lo of the bias of that column (slot NST QV).  One scale 2^gex > max |v| for all of v; returns
// 2^(gex - 46).  red: one double per wavefront of LDS, free until the block's next barrier.
template <int NCG, int NST>
__device__ inline double fz_v_operands(const FzArgs &a, uint4 *qs, double *red)
{
    constexpr int NSLOT = NST * QV + 1;
    const int N = a.N, M = a.M, K = a.K, KM = a.KM;
    const double *v = a.v;
    const int nthr = blockDim.x, nwv = nthr >> 6;
    double mx = 0.0;
    {
        const long long NP = (long long)N + M + (long long)N * M, nv = K * NP;
        for (long long q0 = threadIdx.x; q0 < nv; q0 += 4 * nthr) {
            double x[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const long long q = q0 + (long long)u * nthr;
                x[u] = q < nv ? v[(q / NP) * a.L + q % NP] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++) mx = fmax(mx, fabs(x[u]));
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = red[0];
    for (int q = 1; q < nwv; q++) mx = fmax(mx, red[q]);
    int gex = 0;
    if (mx > 0.0) (void)frexp(mx, &gex);
    const double sc = ldexp(1.0, VBITS - gex);
    uint32_t *o = (uint32_t *)qs;
    // item (cg, st, lane, b): dword b of lane (g, lo) in the slots of spin tile st (st = NST:
    // the bias slot) of column group cg
    for (int item = threadIdx.x; item < NCG * (NST + 1) * 256; item += nthr) {
        const int b = item & 3, lane = (item >> 2) & 63, rest = item >> 8;
        const int st = rest % (NST + 1), cg = rest / (NST + 1);
        const int lo = lane & 15, g = lane >> 4;
        uint32_t dg[QV] = {};
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int col = 64 * cg + 16 * g + 4 * b + d;
            double val = 0.0;
            if (col < KM) {
                const long long base = (long long)(col / M) * a.L;
                if (st < NST) {
                    const int i = 16 * st + lo;
                    val = i < N ? v[base + N + M + (long long)i * M + col % M] : 0.0;
                } else {
                    val = v[base + N + col % M];
                }
            }
            long long x = (long long)rint(val * sc);
#pragma unroll
            for (int u = 0; u < QV; u++) {
                const long long dd = (u < QV - 1) ? bal8l(x) : x;
                dg[u] |= ((uint32_t)dd & 255u) << (8 * d);
                x = (x - dd) >> 8;
            }
        }
        uint32_t *oc = o + (size_t)cg * NSLOT * 256;
        if (st < NST) {
#pragma unroll
            for (int u = 0; u < QV; u++) oc[((st * QV + u) * 64 + lane) * 4 + b] = dg[u];
        } else {
            uint32_t bw = 0u;
#pragma unroll
            for (int u = 0; u < QV; u++)
                if (lo == u) bw = dg[u];
            oc[(NST * QV * 64 + lane) * 4 + b] = bw;
        }
    }
    __syncthreads();
    return mx > 0.0 ? ldexp(1.0, gex - VBITS) : 0.0;
}

// One buffer (two, by super-chunk parity) of the one-pass kernel's LDS, in dwords: the
// super-chunk's digit image XT (copied from Ug as it is), the weight digits, the spin
// masks, the wavefronts' largest |omega|, and the super-chunk's rows' scales (ES, FP64),
// c-column integers (XC, KS per row) and spin words (SP), also copied from global memory.
template <int NCG, int NST, int W, int KS>
struct FzLds {
    static constexpr int NCT = 4 * NCG;
    static constexpr int XT_DW = NCT * QU * (W / 4) * 256;  // [column tile][t][chunk] 64 x 16 bytes
    static constexpr int OD_DW = 8 * 4 * W;  // [chunk][g'][u, 8 slots][row tile of the chunk]
    static constexpr int MK_DW = NST * 16 * 4 * W;  // [st][g'][chunk][i][row tile of the chunk]
    static constexpr int ES_OFF = XT_DW + OD_DW + MK_DW + 2 * W;  // 16-byte aligned
    static constexpr int XC_OFF = ES_OFF + 32 * W;
    static constexpr int SP_OFF = XC_OFF + 16 * W * KS;
    static constexpr int BUF_DW = SP_OFF + 16 * W;
    static constexpr int NSLOT = NST * QV + 1;
    static constexpr int NQ = NCG * NSLOT * 64;  // uint4 of the forward B operands
    static constexpr int XT_U4 = XT_DW / 4;      // uint4 of one super-chunk's digit image
};


// ---- the one-pass S*v (MODE 0: forward and backward; MODE 1: backward with the
// moment weights 1 or Re E_s).  A block of W wavefronts takes super-chunks of 16 W rows
// (W / 4 chunks of 64 rows, the K dimension of the backward products): super-chunks
// [b nsc / nb, (b + 1) nsc / nb), nsc = rpad / (16 W).  Wavefront wv runs the forward
// product of row tile wv and the backward products of column tiles [wv CTW, wv CTW + CTW)
// (CTW = 4 NCG / W) for every spin tile.  Writes part[b][0 .. P] (P: the sum of the
// sample weights).
template <int NCG, int NST, int KS, int MODE>
__global__ __launch_bounds__(64 * fz_waves(NCG), 2) void k_q8_fused(FzArgs a)
{
    constexpr int W = fz_waves(NCG);
    constexpr int CTW = 4 * NCG / W;
    constexpr int NKC = W / 4;  // chunks of 64 rows per super-chunk
    using L = FzLds<NCG, NST, W, KS>;
    constexpr int NSF = QU + QV - 1 - QSKIP;  // digit weights 256^s, s = QSKIP .. QU + QV - 2
    // LDS: the forward B operands and the two buffers, as distinct arrays, so that the
    // compiler sees that the copies into one buffer do not alias the reads of the other
    // (it would otherwise wait for the copies before every LDS read)
    __shared__ __align__(16) uint4 qs[MODE == 0 ? L::NQ : 1];
    __shared__ __align__(16) uint32_t fzb0[L::BUF_DW], fzb1[L::BUF_DW];

    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, lo = lane & 15;
    const int N = a.N, M = a.M, KM = a.KM;
    const long long R = a.R, P = a.P;
    long long c0, c1;
    {
        const long long nsc = a.n64 / NKC, nb = gridDim.x, b = blockIdx.x;
        c0 = nsc * b / nb;
        c1 = nsc * (b + 1) / nb;
    }
    double vsc = 0.0, ov = 0.0;
    double qa[NST][KS];  // v_a[k][16 st + lo] / vsc: the c columns' forward term in digit units
    if constexpr (MODE == 0) {
        vsc = fz_v_operands<NCG, NST>(a, qs, (double *)fzb0);
        ov = *a.ov;
        const double inv = vsc > 0.0 ? 1.0 / vsc : 0.0;  // a power of two
#pragma unroll
        for (int st = 0; st < NST; st++)
#pragma unroll
            for (int k = 0; k < KS; k++) {
                const int i = 16 * st + lo;
                qa[st][k] = i < N ? a.v[(long long)k * a.L + i] * inv : 0.0;
            }
    }
    // the lane's bias digit weight (forward bias slot: lane lo holds digit lo)
    const double pw = lo < QV ? ldexp(1.0, 8 * lo) : 0.0;
    const int j4 = lo & 3;
    // the transpose loads (ds_read_b64_tr_b8: within each 16 lanes, lane x gets byte x % 8
    // of the 8 bytes that lane 2 e + x / 8 read, as its byte e): lane (g, lo) reads the 8
    // bytes of row 16 (lo / 8) + 4 g + (lo / 2) % 4, columns 8 (lo % 2) .. + 7, so that
    // lane (g, x) receives column x of rows 4 g + j and 16 + 4 g + j (j < 4) in order
    const int trofs = (16 * (lo >> 3) + 4 * g + ((lo >> 1) & 3)) * 4 + 2 * (lo & 1);

    v4i acc[CTW][NST][NSF];  // sum_r b_r[i] X Omega by digit weight 256^(s + QSKIP)
    v4i tt[CTW][QU];         // sum_r Omega_u X_t: row u of the result
#pragma unroll
    for (int c = 0; c < CTW; c++) {
#pragma unroll
        for (int st = 0; st < NST; st++)
#pragma unroll
            for (int s = 0; s < NSF; s++) acc[c][st][s] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < QU; t++) tt[c][t] = v4i{0, 0, 0, 0};
    }
    double acc_a[NST][KS];  // sum_r sigma_r[16 st + lo] Xc[r][k] omega_r over the lane group's rows
#pragma unroll
    for (int st = 0; st < NST; st++)
#pragma unroll
        for (int k = 0; k < KS; k++) acc_a[st][k] = 0.0;
    double wsum = 0.0;
    int G = 0;
    bool gvalid = false, dirty = false, first = true;
    double *part = a.part + (long long)blockIdx.x * (P + 1);

    // accumulators -> FP64 into this block's partial row (at scale 2^(G - 46)), cleared
    auto flush = [&]() {
        const double osc = ldexp(1.0, G - WBITS) * 65536.0;  // 256^QSKIP
#pragma unroll
        for (int c = 0; c < CTW; c++) {
            const int ct = wv * CTW + c, col = 16 * ct + lo;
            // T_s of column col: rows u = s - t of the T products (u = 4 g_src + e_src)
            double ts[NSF];
#pragma unroll
            for (int s = 0; s < NSF; s++) {
                double x = 0.0;
#pragma unroll
                for (int t = 0; t < QU; t++) {
                    const int u = s + QSKIP - t;
                    if (u >= 0 && u < QW) x += (double)__shfl(tt[c][t][u & 3], 16 * (u >> 2) + lo);
                }
                ts[s] = x;
            }
            const bool colok = col < KM;
            const long long kb = colok ? (long long)(col / M) * a.L : 0;
            const int jc = colok ? col % M : 0;
#pragma unroll
            for (int st = 0; st < NST; st++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    double hs = 0.0;
#pragma unroll
                    for (int s = NSF - 1; s >= 0; s--) hs = fma(hs, 256.0, fma(-2.0, (double)acc[c][st][s][e], ts[s]));
                    const int i = 16 * st + 4 * g + e;
                    if (colok && i < N) {
                        double *p = part + kb + N + M + (long long)i * M + jc;
                        const double val = hs * osc;
                        *p = first ? val : *p + val;
                    }
                }
            if (g == 0 && colok) {
                double hb = 0.0;
#pragma unroll
                for (int s = NSF - 1; s >= 0; s--) hb = fma(hb, 256.0, ts[s]);
                double *p = part + kb + N + jc;
                const double val = hb * osc;
                *p = first ? val : *p + val;
            }
#pragma unroll
            for (int st = 0; st < NST; st++)
#pragma unroll
                for (int s = 0; s < NSF; s++) acc[c][st][s] = v4i{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < QU; t++) tt[c][t] = v4i{0, 0, 0, 0};
        }
        first = false;
    };

    // Super-chunk ch's digit image, row scales, c-column integers and spin words go from
    // global memory straight into LDS buffer p (global_load_lds_dwordx4: lane l's 16 bytes
    // land at the instruction's LDS base + 16 l), issued one super-chunk ahead; the wavefronts
    // share the copies (the digits W ways, the per-row arrays one wavefront each).
    auto dma = [&](long long ch, uint32_t *b) {
        // the lane offset is formed here (an opaque copy): hoisted out of the loop, the
        // per-lane addresses would hold registers through the whole kernel
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const uint4 *src = a.Ug + ch * L::XT_U4;
        for (int i = wv; i < L::XT_U4 / 64; i += W)
            __builtin_amdgcn_global_load_lds((const void *)(src + i * 64 + ln),
                                             (__attribute__((address_space(3))) void *)(b + i * 256), 16, 0, 0);
        const long long r0 = ch * 16 * W;  // the super-chunk's first row
        if (wv == W - 1 && ln < 8 * W)     // 16 W doubles
            __builtin_amdgcn_global_load_lds((const void *)((const uint4 *)(a.escale + r0) + ln),
                                             (__attribute__((address_space(3))) void *)(b + L::ES_OFF), 16, 0, 0);
        if (wv == W - 2)                   // 16 W KS floats: W KS / 4 instructions of 64 lanes
            for (int i = 0; i < W * KS / 4; i += 4) {
                if (ln < 16 * (W * KS / 4 - i < 4 ? W * KS / 4 - i : 4))
                    __builtin_amdgcn_global_load_lds(
                        (const void *)((const uint4 *)(a.Xc + r0 * KS) + i * 16 + ln),
                        (__attribute__((address_space(3))) void *)(b + L::XC_OFF + i * 64), 16, 0, 0);
            }
        if (wv == W - 3 && ln < 4 * W)     // 16 W words
            __builtin_amdgcn_global_load_lds((const void *)((const uint4 *)(a.spad + r0) + ln),
                                             (__attribute__((address_space(3))) void *)(b + L::SP_OFF), 16, 0, 0);
    };
    // operands of this wavefront's row tile of the super-chunk, from the LDS copies:
    // A operands of lane (g, lo) = row lo of the tile, the 16 columns of column tile
    // 4 cg + g (the order of the v operands)
    uint4 av[NCG][QU];
    uint32_t word[4];
    double es[4];
    float xc[4][KS];
    auto load_tile = [&](const uint32_t *xt) {
        int ln = lane;  // opaque, as in dma
        asm volatile("" : "+v"(ln));
        const int gl = ln >> 4, ll = ln & 15;
        const uint4 *x4 = (const uint4 *)xt;
#pragma unroll
        for (int cg = 0; cg < NCG; cg++)
#pragma unroll
            for (int t = 0; t < QU; t++)
                av[cg][t] = x4[(((4 * cg + gl) * QU + t) * NKC + (wv >> 2)) * 64 + 16 * (wv & 3) + ll];
        const uint32_t *sp = xt + L::SP_OFF;
        const double *esl = (const double *)(xt + L::ES_OFF);
        const float *xcl = (const float *)(xt + L::XC_OFF);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int row = 16 * wv + 4 * gl + e;
            word[e] = sp[row];
            es[e] = esl[row];  // zero on padding rows
#pragma unroll
            for (int k = 0; k < KS; k++) xc[e][k] = xcl[row * KS + k];
        }
    };

    // super-chunk ch in buffer xt; the copies of ch + 1 go to nx
    auto body = [&](long long ch, uint32_t *xt, uint32_t *nx) {
        uint32_t *od = xt + L::XT_DW;
        uint32_t *mk = od + L::OD_DW;
        double *wmx = (double *)(mk + L::MK_DW);
        const long long rt = ch * W + wv;
        // this super-chunk's copies are complete in every wavefront; every wavefront is past
        // its reads of the other buffer (the previous super-chunk), which the next copy fills
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        load_tile(xt);

        // ---- phase A: this wavefront's 16-row tile
        const int er = lo >> 2;  // this lane's row 4 g + er
        const double esr = er == 0 ? es[0] : (er == 1 ? es[1] : (er == 2 ? es[2] : es[3]));
        const long long rrow = rt * 16 + 4 * g + er;
        double ws;  // the sample weight w_s of row rrow
        if constexpr (MODE == 0) {
            double h[4];
            {
                v4i accb[QU];
#pragma unroll
                for (int t = 0; t < QU; t++) accb[t] = v4i{0, 0, 0, 0};
#pragma unroll
                for (int cg = 0; cg < NCG; cg++) {
                    const v4i bb = as_v4i(qs[(cg * L::NSLOT + NST * QV) * 64 + lane]);
#pragma unroll
                    for (int t = 0; t < QU; t++)
                        accb[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(as_v4i(av[cg][t]), bb, accb[t], 0, 0, 0);
                }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    double x = (double)accb[QU - 1][e];
#pragma unroll
                    for (int t = QU - 2; t >= 0; t--) x = fma(x, 256.0, (double)accb[t][e]);
                    h[e] = x * pw;
                }
            }
#pragma unroll
            for (int st = 0; st < NST; st++) {
                v4i af[NSF];
#pragma unroll
                for (int s = 0; s < NSF; s++) af[s] = v4i{0, 0, 0, 0};
#pragma unroll
                for (int cg = 0; cg < NCG; cg++)
#pragma unroll
                    for (int u = 0; u < QV; u++) {
                        const v4i bv = as_v4i(qs[(cg * L::NSLOT + st * QV + u) * 64 + lane]);
#pragma unroll
                        for (int t = 0; t < QU; t++)
                            if (t + u >= QSKIP)
                                af[t + u - QSKIP] =
                                    __builtin_amdgcn_mfma_i32_16x16x64_i8(as_v4i(av[cg][t]), bv, af[t + u - QSKIP], 0, 0, 0);
                    }
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    // digit weights in pairs first: |af| <= 6 2^20 (three digit pairs and two
                    // column groups of 64 products of at most 2^14), so af + 256 af' fits int32
                    double hs = 0.0;
#pragma unroll
                    for (int s = NSF - 2; s >= 0; s -= 2)
                        hs = fma(hs, 65536.0, (double)(int)((uint32_t)af[s][e] + ((uint32_t)af[s + 1][e] << 8)));
                    hs *= 65536.0;  // 256^QSKIP
#pragma unroll
                    for (int k = 0; k < KS; k++) hs = fma((double)xc[e][k], qa[st][k], hs);
                    const bool dn = (word[e] >> (16 * st + lo)) & 1u;
                    h[e] += dn ? -hs : hs;
                }
            }
            const double wr = fz_sum16x4(h, lo) * vsc * esr;  // row 4 g + er
            // the sample's rows: quads KS (er / KS) .. + KS - 1 of this lane group, added in order
            double wq[4];
#pragma unroll
            for (int e = 0; e < 4; e++) wq[e] = fz_shfl(wr, 16 * g + 4 * e + j4);
            const int e0 = KS * (er / KS);
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < KS; q++)
                s += (e0 + q == 0 ? wq[0] : (e0 + q == 1 ? wq[1] : (e0 + q == 2 ? wq[2] : wq[3])));
            ws = s - ov;
        } else {
            const long long smp = rrow / KS;
            double wt = 1.0;
            if (a.wsel == 2) wt = smp < a.S ? a.E[smp].re : 0.0;
            ws = rrow < R ? wt : 0.0;
        }
        const double om = ws * esr;  // omega of row rrow
        if (j4 == 0 && er % KS == 0 && rrow < R) wsum += ws;

        // a block: acc_a[st][k] += sum_e sigma_{4 g + e}[16 st + lo] Xc[e][k] omega_e
        {
            double omq[4];
#pragma unroll
            for (int e = 0; e < 4; e++) omq[e] = fz_shfl(om, 16 * g + 4 * e);
#pragma unroll
            for (int st = 0; st < NST; st++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const bool dn = (word[e] >> (16 * st + lo)) & 1u;
                    const double z = dn ? -omq[e] : omq[e];
#pragma unroll
                    for (int k = 0; k < KS; k++) acc_a[st][k] = fma((double)xc[e][k], z, acc_a[st][k]);
                }
        }

        // spin masks: spin 16 st + lo, rows 4 g + j
#pragma unroll
        for (int st = 0; st < NST; st++) {
            uint32_t bits = 0u;
#pragma unroll
            for (int j = 0; j < 4; j++) bits |= ((word[j] >> (16 * st + lo)) & 1u) << j;
            mk[((((st * 4 + g) * NKC + (wv >> 2)) * 16 + lo) * 4) + (wv & 3)] = fz_expand4(bits) * 0xffu;
        }
        // weight digits at the block's current scale (rows that do not fit are redone below)
        // (balanced digits: q + 128 (1 + 256 + ... + 256^5) has the plain base-256 digits
        // d_u + 128, so byte u of that sum with its top bit flipped is d_u)
        auto put_digits = [&]() {
            long long q = 0;
            if (gvalid && fabs(om) < ldexp(1.0, G)) q = (long long)rint(om * ldexp(1.0, WBITS - G));
            const unsigned long long b = (unsigned long long)(q + 0x808080808080LL) ^ 0x808080808080ULL;
            if (j4 == 0) {
                uint8_t *ob = (uint8_t *)od + (((wv >> 2) * 4 + g) * 8 * 4 + (wv & 3)) * 4 + er;
#pragma unroll
                for (int u = 0; u < QW; u++) ob[u * 16] = (uint8_t)(b >> (8 * u));
            }
        };
        put_digits();
        {
            double m = fabs(om);
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) m = fmax(m, __shfl_xor(m, off));
            if (lane == 0) wmx[wv] = m;
        }
        __syncthreads();
        {
            double m = wmx[0];
#pragma unroll
            for (int q = 1; q < W; q++) m = fmax(m, wmx[q]);
            int em = 0;
            if (m > 0.0) (void)frexp(m, &em);  // m < 2^em
            if (m > 0.0 && (!gvalid || em > G)) {  // uniform over the block
                if (dirty) {
                    flush();
                    dirty = false;
                }
                G = em + 2;
                gvalid = true;
                put_digits();
                __syncthreads();
            }
        }

        // the next super-chunk's copies run during phase C
        if (ch + 1 < c1) dma(ch + 1, nx);

        // ---- phase C: the super-chunk's backward products for column tiles wv CTW + c
        const uint4 *od4 = (const uint4 *)od;
        const uint4 *mk4 = (const uint4 *)mk;
#pragma unroll
        for (int kc = 0; kc < NKC; kc++) {
            // B operand of lane (g, lo): bytes 4 q + j = row 16 q + 4 g + j of the chunk,
            // column lo: two transpose loads (q = 0, 1 and q = 2, 3) of the rows' 8-byte halves
            uint4 bv[CTW][QU];
#pragma unroll
            for (int c = 0; c < CTW; c++) {
                const int ct = wv * CTW + c;
#pragma unroll
                for (int t = 0; t < QU; t++) {
                    const uint32_t *base = xt + ((ct * QU + t) * NKC + kc) * 256 + trofs;
                    const v2i x0 = lds_tr8(base), x1 = lds_tr8(base + 32 * 4);
                    bv[c][t] = uint4{(uint32_t)x0.x, (uint32_t)x0.y, (uint32_t)x1.x, (uint32_t)x1.y};
                }
            }
            {
                const uint4 ta = od4[(kc * 4 + g) * 8 + (lo < QW ? lo : 0)];
                const v4i tv = lo < QW ? as_v4i(ta) : v4i{0, 0, 0, 0};
#pragma unroll
                for (int c = 0; c < CTW; c++)
#pragma unroll
                    for (int t = 0; t < QU; t++)
                        tt[c][t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(tv, as_v4i(bv[c][t]), tt[c][t], 0, 0, 0);
            }
#pragma unroll
            for (int st = 0; st < NST; st++) {
                const uint4 mm = mk4[((st * 4 + g) * NKC + kc) * 16 + lo];
#pragma unroll
                for (int u = 0; u < QW; u++) {
                    const uint4 o = od4[(kc * 4 + g) * 8 + u];
                    const v4i av8 = v4i{(int)(o.x & mm.x), (int)(o.y & mm.y), (int)(o.z & mm.z), (int)(o.w & mm.w)};
#pragma unroll
                    for (int c = 0; c < CTW; c++)
#pragma unroll
                        for (int t = 0; t < QU; t++)
                            if (t + u >= QSKIP)
                                acc[c][st][t + u - QSKIP] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    av8, as_v4i(bv[c][t]), acc[c][st][t + u - QSKIP], 0, 0, 0);
                }
            }
        }
        if (gvalid) dirty = true;
    };
    if (c0 < c1) dma(c0, fzb0);
    for (long long ch = c0; ch < c1; ch += 2) {
        body(ch, fzb0, fzb1);
        if (ch + 1 < c1) body(ch + 1, fzb1, fzb0);
    }
    if (!gvalid) G = 0;
    flush();

    // a block and the weight sum: lane groups, then wavefronts in order
#pragma unroll
    for (int st = 0; st < NST; st++)
#pragma unroll
        for (int k = 0; k < KS; k++) {
            double x = acc_a[st][k];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            acc_a[st][k] = x;
        }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) wsum += __shfl_xor(wsum, off);
    __syncthreads();  // every wavefront is past its last reads of the buffers
    double *red = (double *)fzb0;
    if (g == 0) {
#pragma unroll
        for (int st = 0; st < NST; st++)
#pragma unroll
            for (int k = 0; k < KS; k++) red[wv * FZ_RED + (st * KS + k) * 16 + lo] = acc_a[st][k];
    }
    if (lane == 0) red[W * FZ_RED + wv] = wsum;
    __syncthreads();
    if (wv == 0) {
        for (int q = lane; q < NST * KS * 16; q += 64) {
            const int st = q / (KS * 16), k = (q / 16) % KS, i = 16 * st + (q & 15);
            double x = red[q];
            for (int w = 1; w < W; w++) x += red[w * FZ_RED + q];
            if (i < N) part[(long long)k * a.L + i] = x;
        }
        if (lane == 0) {
            double x = red[W * FZ_RED];
            for (int w = 1; w < W; w++) x += red[W * FZ_RED + w];
            part[P] = x;
        }
    }
}

template <int NCG, int NST, int MODE>
const void *fz_kernel() { return (const void *)k_q8_fused<NCG, NST, NQES_FZ_K, MODE>; }


template <int MODE>
const void *fz_select(const Q8Shape &sh)
{
    switch (sh.fz_ncg * 8 + sh.fz_nst) {
    case 9: return fz_kernel<1, 1, MODE>();
    case 17: return fz_kernel<2, 1, MODE>();
    case 10: return fz_kernel<1, 2, MODE>();
    case 18: return fz_kernel<2, 2, MODE>();
    default: return nullptr;
    }
}

template <int MODE>
hipError_t fz_allow_lds(const Q8Shape &sh)
{
    const void *fn = fz_select<MODE>(sh);
    return fn ? hipSuccess : hipErrorInvalidValue;  // static LDS only (142 KB at most)
}

template <int MODE>
hipError_t fz_launch(const FzArgs &f, const Q8Shape &sh, hipStream_t st)
{
    const size_t lds = 0;
    const hipError_t e = fz_allow_lds<MODE>(sh);
    if (e != hipSuccess) return e;
    const dim3 grid(sh.fz_blocks), blk(64 * fz_waves(sh.fz_ncg));
    switch (sh.fz_ncg * 8 + sh.fz_nst) {
    case 9: hipLaunchKernelGGL((k_q8_fused<1, 1, NQES_FZ_K, MODE>), grid, blk, lds, st, f); break;
    case 17: hipLaunchKernelGGL((k_q8_fused<2, 1, NQES_FZ_K, MODE>), grid, blk, lds, st, f); break;
    case 10: hipLaunchKernelGGL((k_q8_fused<1, 2, NQES_FZ_K, MODE>), grid, blk, lds, st, f); break;
    default: hipLaunchKernelGGL((k_q8_fused<2, 2, NQES_FZ_K, MODE>), grid, blk, lds, st, f); break;
    }
    return hipGetLastError();
}

FzArgs fz_args(const SvfArgs &a, const Q8Shape &sh)
{
    FzArgs f{};
    f.N = a.N;
    f.M = a.M;
    f.K = a.K;
    f.L = (int)a.L;
    f.KM = a.K * a.M;
    f.S = a.S;
    f.R = a.S * a.K;
    f.P = a.P;
    f.n64 = sh.rpad / 64;
    f.spins = a.spins;
    f.Ug = (const uint4 *)a.Uq;
    f.Xc = (const float *)a.q8m;
    f.spad = (const uint32_t *)((const uint8_t *)a.q8m + (size_t)sh.rpad * a.K * sizeof(float));
    f.escale = a.escale;
    f.v = (const double *)a.v;
    f.ov = (const double *)a.ov;
    f.E = (const nq64::cd *)a.E;
    f.part = (double *)a.part;
    f.wsel = 0;
    return f;
}

}  // namespace

#define NQ_FZ_CAT2(a, b) a##b
#define NQ_FZ_CAT(a, b) NQ_FZ_CAT2(a, b)

// quantise the real factor rows into Ug / Xc / escale (Uq, q8m, escale of the args), and
// let both fused forms of this shape use their LDS (set here, outside any graph capture)
hipError_t NQ_FZ_CAT(nq_q8f_pack_k, NQES_FZ_K)(const SvfArgs &a, const Q8Shape &sh, hipStream_t st)
{
    const long long R = a.S * a.K;
    const long long ntile = sh.rpad / 16;
    long long blocks = (ntile + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    const size_t lds = (size_t)4 * 16 * (a.Ucols + 1) * 4;
    hipError_t e = nq_allow_dynamic_lds((const void *)k_q8_packg<0>, lds);
    if (e == hipSuccess) e = fz_allow_lds<0>(sh);
    if (e == hipSuccess) e = fz_allow_lds<1>(sh);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_q8_packg<0>), dim3((unsigned)blocks), dim3(256), lds, st, (const float *)a.U, R, ntile,
                       a.Ucols, a.K * a.M, a.K, sh.fz_ncg, (uint4 *)a.Uq, (float *)a.q8m, (double *)a.escale, a.Umax,
                       a.spins, (uint32_t *)((uint8_t *)a.q8m + (size_t)sh.rpad * a.K * sizeof(float)));
    return hipGetLastError();
}

// one S*v: out[0 .. P] = the raw sums (sum_s O_s w_s and sum_s w_s)
hipError_t NQ_FZ_CAT(nq_q8f_apply_k, NQES_FZ_K)(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out,
                                                 hipStream_t st)
{
    FzArgs f = fz_args(a, sh);
    hipError_t e = fz_launch<0>(f, sh, st);
    if (e == hipSuccess) e = nq_chunk_reduce(false, a.part, a.P + 1, a.P + 1, sh.fz_blocks, scratch, out, st);
    return e;
}

// backward sums with the moment weights (wsel 1: ones, 2: Re E_s) into part, reduced into out
hipError_t NQ_FZ_CAT(nq_q8f_moment_k, NQES_FZ_K)(const SvfArgs &a, const Q8Shape &sh, int wsel, void *part,
                                                  void *scratch, void *out, hipStream_t st)
{
    FzArgs f = fz_args(a, sh);
    f.part = (double *)part;
    f.wsel = wsel;
    hipError_t e = fz_launch<1>(f, sh, st);
    if (e == hipSuccess) e = nq_chunk_reduce(false, part, a.P + 1, a.P + 1, sh.fz_blocks, scratch, out, st);
    return e;
}
