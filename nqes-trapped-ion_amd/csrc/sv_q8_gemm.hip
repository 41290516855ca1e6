// sv_q8_gemm.hip -- forward pass of the int8 S*v (real weights, N <= 32) as one GEMM
// over the columns: P_ri = sum_c X_rc Q_ci on v_mfma_i32_16x16x64_i8, then the signs of
// the spins and the digit weights per row.  The factor digits come in a row-major copy
// (Uf, written by the pack kernel of sv_q8.hip next to its tiles): per 16-row tile and
// digit plane t, 16 qcols bytes, chunk 16 ct + l (l < 16) = the 16 digits of row l of
// the tile at columns 16 ct .. 16 ct + 15.  When the K c columns would open a column
// tile of their own (K M a multiple of 16, K in {1, 2, 4}), the digits cover the K M
// tanh columns only and the c columns' integers (xc) enter in FP64.  Compiled with MFMA results in VGPRs (the
// sums are read by VALU code once per tile).  Reference: the factored S*v of
// rbm.cc:661-684, as in sv_q8.hip.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <type_traits>

#include "kernels.hpp"
#include "sv_q8_common.hpp"

namespace {

// DPP move of a double: lane l takes the value of its partner under CTRL
template <int CTRL>
__device__ inline double dpp_mov(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The four sums h[e] over the 16 lanes of a row at once: a row mirror splits them in
// pairs (lanes 0-7 keep e = 0, 1, lanes 8-15 e = 2, 3), a half-row mirror in singles,
// two quad swaps finish them.  Lane l returns the sum of h[l / 4] (every lane of the
// row contributes exactly once).
__device__ inline double dpp_sum16x4(const double (&h)[4], int l)
{
    const bool lowhalf = l < 8, lowquad = (l & 7) < 4;
    const double k0 = lowhalf ? h[0] : h[2], k1 = lowhalf ? h[1] : h[3];
    const double s0 = lowhalf ? h[2] : h[0], s1 = lowhalf ? h[3] : h[1];
    const double a0 = k0 + dpp_mov<0x140>(s0), a1 = k1 + dpp_mov<0x140>(s1);
    double b = (lowquad ? a0 : a1) + dpp_mov<0x141>(lowquad ? a1 : a0);
    b += dpp_mov<0x4E>(b);
    b += dpp_mov<0xB1>(b);
    return b;
}

// ---- GEMM forward pass (N <= 32, at most 256 digit columns).
//   w_r = (1 / f_r) 2^(g - 46) sum_c X_rc (sum_i sigma_ri Q_ci + bq_c)
// with one scale 2^g > max |v| for all of v and its bias (absolute error 2^-47 of the
// largest entry, the size of FP64 rounding in a dot product).  The sum over columns
// comes first, on int8 MFMAs of 64 columns, P_ri = sum_c X_rc Q_ci for 16 rows x 16
// spins per spin tile and digit pair (t, u); the signs after it.  The bias takes one
// more tile per column group whose spin slot u holds digit u of bq, so that its 18 digit
// pairs cost 3 MFMAs.  Operands: A lane l, byte e = digit of row 16 rt + l%16 at column
// 64 cg + 16 (l/16) + e (the Uf chunk 64 cg + l); B lane l, byte e = digit of the same
// column for spin 16 st + l%16 (slot st QV + u), or digit l%16 of bq (slot NST QV).
// Qg: [cg][slot][lane] uint4, then the scale 2^(g - 46) as a double; with the c columns
// apart (a.qcols < a.Ucols) then qa[k 32 + i] = v_a[k][i] / 2^(g - 46) (FP64, unquantised;
// 0 for i >= N), the c columns' B values, 16 bytes after the scale.  blockIdx.z: the set
// (v + z vss doubles, qg + z qss uint4) when several vectors are prepared at once.
__global__ __launch_bounds__(256) void k_q8_prep_vg(SvfArgs a, int ncg, int nst, uint4 *qg, long long vss,
                                                    long long qss)
{
    __shared__ double red[4];
    const int N = a.N, M = a.M, K = a.K, QC = a.qcols;
    const long long KM = (long long)K * M;
    const double *v = (const double *)a.v + blockIdx.z * vss;
    qg += blockIdx.z * qss;
    auto Bval = [&](int col, int i) -> double {
        if (col < KM) return v[(long long)(col / M) * a.L + N + M + (long long)i * M + col % M];
        if (col < KM + K) return v[(long long)(col - KM) * a.L + i];
        return 0.0;
    };
    auto bias = [&](int col) -> double { return col < KM ? v[(long long)(col / M) * a.L + N + col % M] : 0.0; };
    // maximum over v (the parameters of the K chains, L apart, N + M + N M each); every
    // block takes it, loads four ahead
    double mx = 0.0;
    {
        const long long NP = (long long)N + M + (long long)N * M, nv = K * NP, step = blockDim.x;
        for (long long q0 = threadIdx.x; q0 < nv; q0 += 4 * step) {
            double x[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const long long q = q0 + u * step;
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
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    int gex = 0;
    if (mx > 0.0) (void)frexp(mx, &gex);
    const double sc = ldexp(1.0, VBITS - gex);
    // block (cg, st), st = nst: the bias; wavefront w forms dword w of every lane's uint4
    // (its columns e = 4 w .. 4 w + 3 of the lane's 16)
    const int lane = threadIdx.x & 63, lo = lane & 15, w = threadIdx.x >> 6;
    const int cg = blockIdx.x / (nst + 1), st = blockIdx.x % (nst + 1);
    const int nslot = nst * QV + 1;
    uint32_t dg[QV] = {};
#pragma unroll
    for (int e4 = 0; e4 < 4; e4++) {
        const int col = 64 * cg + 16 * (lane >> 4) + 4 * w + e4;
        double val = 0.0;
        if (col < QC) {
            if (st < nst) val = 16 * st + lo < N ? Bval(col, 16 * st + lo) : 0.0;
            else val = bias(col);
        }
        long long x = (long long)rint(val * sc);
#pragma unroll
        for (int u = 0; u < QV; u++) {
            const long long d = (u < QV - 1) ? bal8l(x) : x;
            dg[u] |= ((uint32_t)d & 255u) << (8 * e4);
            x = (x - d) >> 8;
        }
    }
    uint32_t *o = (uint32_t *)(qg + (size_t)cg * nslot * 64);
    if (st < nst) {
#pragma unroll
        for (int u = 0; u < QV; u++) o[((st * QV + u) * 64 + lane) * 4 + w] = dg[u];
    } else {
        // slot nst QV, lane (g, lo): digit lo of the bias columns of lane group g (every
        // lane of the group computed the same bias values)
        uint32_t bw = 0u;
#pragma unroll
        for (int u = 0; u < QV; u++)
            if (lo == u) bw = dg[u];
        o[(nst * QV * 64 + lane) * 4 + w] = bw;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *(double *)(qg + (size_t)ncg * nslot * 64) = mx > 0.0 ? ldexp(1.0, gex - VBITS) : 0.0;
    if (blockIdx.x == 0 && QC < a.Ucols && (int)threadIdx.x < 32 * K) {
        const int k = threadIdx.x >> 5, i = threadIdx.x & 31;
        ((double *)(qg + (size_t)ncg * nslot * 64 + 1))[threadIdx.x] = i < N ? v[(long long)k * a.L + i] * sc : 0.0;
    }
}

// GEMM forward pass.  Wavefronts loop over 16-row tiles (NCG (NST QV + 1) QU MFMAs
// into 8 int32 accumulators per spin tile, digit weights 256^(t + u), and 3 for the
// bias).  Per row: each lane
// sums its spin (FP64 Horner over the digit weights, sign of its spin) and bias digit,
// then 16 lanes by DPP.  Output wrow[r].
// KW = K in 1, 2, 4: the sample weights w_s written directly (no combine step); KW = 0:
// the row sums wrow[r]
// blockIdx.z: the digit set (a.Uf + z fss bytes, a.vq + z vss bytes, a.w + z wss
// doubles, a.xc + z xss floats) when several sets run in one launch (KW != 0 only: the
// row sums of KW = 0 share one buffer)
// CS (KW != 0): the c columns apart, a.qcols = K M digit columns; each lane adds
// sigma_ri sum_k X_c[r][k] qa[k][i] (integers times the unquantised v_a in units of the v
// scale, FP64) to its rows after the last spin tile's sums.  A tile's 16 KW X_c values
// load into one register per lane with the factor digits and pass through the
// wavefront's 256 bytes of LDS (registers across the MFMA phases would spill).
template <int NST, int NCG, int KW, bool CS>
__global__ __launch_bounds__(256, NST * NCG > 6 ? 2 : 3) void k_q8_fwdg(SvfArgs a, long long fss, long long vss,
                                                                         long long wss, long long xss)
{
    a.Uf = (const uint8_t *)a.Uf + blockIdx.z * fss;
    a.vq = (uint8_t *)a.vq + blockIdx.z * vss;
    a.w = (double *)a.w + blockIdx.z * wss;
    a.xc = a.xc + blockIdx.z * xss;
    constexpr int NSF = QU + QV - 1;
    constexpr int NSLOT = NST * QV + 1;
    constexpr int NQ = NCG * NSLOT * 64;
    constexpr int NQL = CS ? NQ + 1 + 16 * KW : NQ;  // uint4 in LDS: digits (, scale, qa)
    extern __shared__ __align__(16) unsigned char smem_g[];
    uint4 *qs = (uint4 *)smem_g;
    {
        const uint4 *src = (const uint4 *)a.vq;
        for (int q = threadIdx.x; q < NQL; q += blockDim.x) qs[q] = src[q];
    }
    const double vsc = *(const double *)((const uint4 *)a.vq + NQ);
    __syncthreads();
    const double *qa = (const double *)(qs + NQ + 1);
    __shared__ __align__(16) float xlds[4][64];  // per wavefront: a tile's X_c, row-major [16][KW]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, lo = lane & 15;
    const int NW = a.NW, NCT = a.qcols / 16;
    const long long R = a.S * a.K;
    const long long ntile = (R + 15) / 16;
    const uint4 *Uf = (const uint4 *)a.Uf;
    double *wrow = (double *)a.wrow, *wout = (double *)a.w;
    const double ov = KW ? *(const double *)a.ov : 0.0;
    const double pw = lo < QV ? ldexp(1.0, 8 * lo) : 0.0;  // weight of this lane's bias digit
    const long long step = (long long)gridDim.x * (blockDim.x >> 6);
    // Loads are unconditional (no branch between a load and its use, so the waits count
    // only older loads): lanes past the last column read chunk 16 NCT - 1, whose products
    // meet zero v digits; rows past R read row R - 1 (their sums are not written).  Three
    // wavefronts per SIMD (two for NST NCG = 8, whose 168 registers would spill) cover
    // each other's load latency (measured 5 % faster than two with the next tile's
    // digits loaded into registers ahead).
    const int nq16 = 16 * NCT;
    for (long long rt = (long long)blockIdx.x * (blockDim.x >> 6) + wv; rt < ntile; rt += step) {
        uint4 av[NCG][QU];
        uint32_t word[4];
        double es[4];
#pragma unroll
        for (int cg = 0; cg < NCG; cg++) {
            const int q = 64 * cg + lane < nq16 ? 64 * cg + lane : nq16 - 1;
#pragma unroll
            for (int t = 0; t < QU; t++) av[cg][t] = Uf[(rt * QU + t) * nq16 + q];
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const long long r = rt * 16 + 4 * g + e, rc = r < R ? r : R - 1;
            word[e] = a.spins[rc * NW];
            es[e] = a.escale[rc];
        }
        // the tile's X_c (rows up to the padded count hold zeros past R); lanes >= 16 KW
        // read a duplicate
        const float xl = CS ? a.xc[rt * 16 * KW + (lane & (16 * KW - 1))] : 0.f;
        // an opaque lane offset keeps the v operand reads inside the loop (hoisted, they
        // would hold NCG NSLOT 4 registers)
        int ql = lane;
        asm volatile("" : "+v"(ql));
        // Phases: the bias tile, then each spin tile, the FP64 sums of one phase written
        // after the MFMAs of the next (independent of them: they issue while those run;
        // measured 6% faster than all sums after all MFMAs).
        // Lane (g, lo), register e: row 16 rt + 4 g + e, spin 16 st + lo / bias digit lo.
        double h[4];
        v4i accb[QU];
#pragma unroll
        for (int t = 0; t < QU; t++) accb[t] = v4i{0, 0, 0, 0};
#pragma unroll
        for (int cg = 0; cg < NCG; cg++) {
            v4i at[QU], bb;
#pragma unroll
            for (int t = 0; t < QU; t++) __builtin_memcpy(&at[t], &av[cg][t], 16);
            const uint4 b4 = qs[(cg * NSLOT + NST * QV) * 64 + ql];
            __builtin_memcpy(&bb, &b4, 16);
#pragma unroll
            for (int t = 0; t < QU; t++) accb[t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(at[t], bb, accb[t], 0, 0, 0);
        }
        if constexpr (CS) {
            if (lane < 16 * KW) xlds[wv][lane] = xl;
            __builtin_amdgcn_wave_barrier();
        }
        v4i accp[NSF];  // sums of the previous spin tile
#pragma unroll
        for (int st = 0; st <= NST; st++) {
            v4i acc[NSF];
            if (st < NST) {
#pragma unroll
                for (int s = 0; s < NSF; s++) acc[s] = v4i{0, 0, 0, 0};
#pragma unroll
                for (int cg = 0; cg < NCG; cg++) {
                    v4i at[QU];
#pragma unroll
                    for (int t = 0; t < QU; t++) __builtin_memcpy(&at[t], &av[cg][t], 16);
#pragma unroll
                    for (int u = 0; u < QV; u++) {
                        v4i bv;
                        const uint4 b4 = qs[(cg * NSLOT + st * QV + u) * 64 + ql];
                        __builtin_memcpy(&bv, &b4, 16);
#pragma unroll
                        for (int t = 0; t < QU; t++)
                            if (t + u >= QSKIP)
                                acc[t + u] = __builtin_amdgcn_mfma_i32_16x16x64_i8(at[t], bv, acc[t + u], 0, 0, 0);
                    }
                }
            }
            if (st == 0) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    double x = (double)accb[QU - 1][e];
#pragma unroll
                    for (int t = QU - 2; t >= 0; t--) x = fma(x, 256.0, (double)accb[t][e]);
                    h[e] = x * pw;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    double hs = (double)accp[NSF - 1][e];
#pragma unroll
                    for (int s = NSF - 2; s >= 0; s--) hs = fma(hs, 256.0, (double)accp[s][e]);
                    const bool dn = (word[e] >> (16 * (st - 1) + lo)) & 1u;
                    h[e] += dn ? -hs : hs;
                }
            }
            if (st < NST) {
#pragma unroll
                for (int s = 0; s < NSF; s++) accp[s] = acc[s];
            }
        }
        if constexpr (CS) {
            // the c columns: row 4 g + e, column k at xlds[wv][(4 g + e) KW + k]
            float xr[4 * KW];
            const float4 *xp = (const float4 *)(xlds[wv] + 4 * g * KW);
#pragma unroll
            for (int q = 0; q < KW; q++) {
                const float4 x4 = xp[q];
                xr[4 * q] = x4.x;
                xr[4 * q + 1] = x4.y;
                xr[4 * q + 2] = x4.z;
                xr[4 * q + 3] = x4.w;
            }
#pragma unroll
            for (int st = 0; st < NST; st++) {
                double qk[KW];
#pragma unroll
                for (int k = 0; k < KW; k++) qk[k] = qa[k * 32 + 16 * st + (ql & 15)];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    double c = 0.0;
#pragma unroll
                    for (int k = 0; k < KW; k++) c = fma((double)xr[e * KW + k], qk[k], c);
                    const bool dn = (word[e] >> (16 * st + lo)) & 1u;
                    h[e] += dn ? -c : c;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        {
            const double hr = dpp_sum16x4(h, lo);  // row e = lo / 4
            const int e = lo >> 2;
            const double ese = e == 0 ? es[0] : (e == 1 ? es[1] : (e == 2 ? es[2] : es[3]));
            const long long r = rt * 16 + 4 * g + e;
            const double wr = hr * vsc * ese;  // the row sum, at lanes l % 4 == 0
            if (KW == 0) {
                if ((lo & 3) == 0 && r < R) wrow[r] = wr;
            } else {
                // w_s = sum over the sample's KW rows (lanes 4 apart, added in order as the
                // combine kernel does) - o^T v, at the lane of its first row
                double acc = 0.0 + wr;
                if (KW >= 2) acc += dpp_mov<0x104>(wr);  // row shift left by 4: lane l + 4
                if (KW == 4) {
                    acc += dpp_mov<0x108>(wr);
                    acc += dpp_mov<0x10C>(wr);
                }
                if ((lo & (4 * KW - 1)) == 0 && r < R) wout[r / KW] = acc - ov;
            }
        }
    }
}

}  // namespace

// the v operands of nsets vectors (a.v + z vss) into a.vq + z qss_bytes, one launch
hipError_t nq_q8_prep_vg_sets(const SvfArgs &a, const Q8Shape &sh, int nsets, long long vss, size_t qss_bytes,
                              hipStream_t st)
{
    hipLaunchKernelGGL(k_q8_prep_vg, dim3(sh.ncg64 * (sh.nst + 1), 1, nsets), dim3(256), 0, st, a, sh.ncg64, sh.nst,
                       (uint4 *)a.vq, vss, (long long)(qss_bytes / 16));
    return hipGetLastError();
}

hipError_t nq_q8_fwd_gemm(const SvfArgs &a, const Q8Shape &sh, hipStream_t st, bool rows, bool prep, int nsets,
                          long long fss, long long vss, long long wss, long long xss)
{
    const bool cs = a.qcols < a.Ucols;
    if (nsets > 1 && (rows || !(a.K == 1 || a.K == 2 || a.K == 4))) return hipErrorInvalidValue;
    if (cs && (rows || !(a.K == 1 || a.K == 2 || a.K == 4) || !a.xc)) return hipErrorInvalidValue;
    if (prep)
        hipLaunchKernelGGL(k_q8_prep_vg, dim3(sh.ncg64 * (sh.nst + 1)), dim3(256), 0, st, a, sh.ncg64, sh.nst,
                           (uint4 *)a.vq, 0LL, 0LL);
    const size_t lds = cs ? sh.qg_bytes : sh.qg_bytes - 16;
    const dim3 grid(sh.fwd_blocks, 1, nsets), blk(256);
#define NQ_FWDK(NST, NCG, KW)                                                                                       \
    do {                                                                                                           \
        if (cs) hipLaunchKernelGGL((k_q8_fwdg<NST, NCG, KW, true>), grid, blk, lds, st, a, fss, vss, wss, xss);   \
        else hipLaunchKernelGGL((k_q8_fwdg<NST, NCG, KW, false>), grid, blk, lds, st, a, fss, vss, wss, xss);     \
    } while (0)
#define NQ_FWDG(NST, NCG)                                                                                              \
    do {                                                                                                              \
        if (rows) hipLaunchKernelGGL((k_q8_fwdg<NST, NCG, 0, false>), grid, blk, lds, st, a, fss, vss, wss, xss);    \
        else if (a.K == 1) NQ_FWDK(NST, NCG, 1);                                                                      \
        else if (a.K == 2) NQ_FWDK(NST, NCG, 2);                                                                      \
        else if (a.K == 4) NQ_FWDK(NST, NCG, 4);                                                                      \
        else hipLaunchKernelGGL((k_q8_fwdg<NST, NCG, 0, false>), grid, blk, lds, st, a, fss, vss, wss, xss);         \
    } while (0)
    switch (sh.nst * 8 + sh.ncg64) {
    case 9: NQ_FWDG(1, 1); break;
    case 10: NQ_FWDG(1, 2); break;
    case 11: NQ_FWDG(1, 3); break;
    case 12: NQ_FWDG(1, 4); break;
    case 17: NQ_FWDG(2, 1); break;
    case 18: NQ_FWDG(2, 2); break;
    case 19: NQ_FWDG(2, 3); break;
    default: NQ_FWDG(2, 4); break;
    }
#undef NQ_FWDG
#undef NQ_FWDK
    return hipGetLastError();
}
