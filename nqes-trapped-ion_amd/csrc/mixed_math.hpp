// mixed_math.hpp -- scalar types of the mixed-precision device path.
//
// T = float (real weights) or cf (complex weights): per-hidden-unit quantities
//     (tanh theta, cosh/sinh tables, O) in FP32.
// D = double or nq64::cd: per-chain K x K matrices (Phi, its inverse, H_loc),
//     amplitude ratios and all sums, in FP64.
// The real instantiation is used when every weight has a zero imaginary part
// (then Phi, H_loc, O and the SR solution stay real, as in the complex algorithm).
#pragma once

#include <hip/hip_runtime.h>

#include "cplx64.hpp"

namespace nqm {

using nq64::cd;

struct cf {
    float re, im;
};

// per-element (T) and accumulation (D) types of the real / complex instantiation
template <bool CPLX> struct Ty;
template <> struct Ty<false> {
    using T = float;
    using D = double;
};
template <> struct Ty<true> {
    using T = cf;
    using D = cd;
};

// ---------------------------------------------------------------- float / cf
__device__ __host__ inline cf mkf(float re, float im)
{
    cf z;
    z.re = re;
    z.im = im;
    return z;
}
__device__ inline float addT(float a, float b) { return a + b; }
__device__ inline cf addT(cf a, cf b) { return mkf(a.re + b.re, a.im + b.im); }
__device__ inline float subT(float a, float b) { return a - b; }
__device__ inline cf subT(cf a, cf b) { return mkf(a.re - b.re, a.im - b.im); }
__device__ inline float negT(float a) { return -a; }
__device__ inline cf negT(cf a) { return mkf(-a.re, -a.im); }
__device__ inline float mulT(float a, float b) { return a * b; }
__device__ inline cf mulT(cf a, cf b) { return mkf(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ inline float divT(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ inline cf divT(cf a, cf b)
{
    float r = __builtin_amdgcn_rcpf(b.re * b.re + b.im * b.im);
    return mkf((a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r);
}
// a + s b for a sign s = +-1: one FMA per component, equal bit for bit to a + b or a - b
__device__ inline float addsT(float a, float s, float b) { return fmaf(s, b, a); }
__device__ inline cf addsT(cf a, float s, cf b) { return mkf(fmaf(s, b.re, a.re), fmaf(s, b.im, a.im)); }
__device__ inline float scaleT(float s, float a) { return s * a; }
__device__ inline cf scaleT(float s, cf a) { return mkf(s * a.re, s * a.im); }
template <typename T> __device__ inline T oneT();
template <> __device__ inline float oneT<float>() { return 1.0f; }
template <> __device__ inline cf oneT<cf>() { return mkf(1.0f, 0.0f); }
template <typename T> __device__ inline T zeroT();
template <> __device__ inline float zeroT<float>() { return 0.0f; }
template <> __device__ inline cf zeroT<cf>() { return mkf(0.0f, 0.0f); }

// tanh and cosh of theta in FP32 (fast device intrinsics; |x| clamped so that
// exp does not overflow -- tanh saturates long before)
__device__ inline float tanhT(float x) { return tanhf(x); }
__device__ inline cf tanhT(cf z)
{
    float sy, cy;
    __sincosf(z.im, &sy, &cy);
    float x = fminf(fmaxf(z.re, -20.0f), 20.0f);
    float e = __expf(x), ie = __expf(-x);
    float sh = 0.5f * (e - ie), ch = 0.5f * (e + ie);
    float d = __builtin_amdgcn_rcpf(sh * sh + cy * cy);
    return mkf(sh * ch * d, sy * cy * d);
}

// ---------------------------------------------------------------- double / cd
__device__ inline double addD(double a, double b) { return a + b; }
__device__ inline cd addD(cd a, cd b) { return nq64::add(a, b); }
__device__ inline double subD(double a, double b) { return a - b; }
__device__ inline cd subD(cd a, cd b) { return nq64::sub(a, b); }
__device__ inline double negD(double a) { return -a; }
__device__ inline cd negD(cd a) { return nq64::neg(a); }
__device__ inline double mulD(double a, double b) { return a * b; }
__device__ inline cd mulD(cd a, cd b) { return nq64::mul(a, b); }
__device__ inline double scaleD(double s, double a) { return s * a; }
__device__ inline cd scaleD(double s, cd a) { return nq64::scale(s, a); }
__device__ inline double conjD(double a) { return a; }
__device__ inline cd conjD(cd a) { return nq64::conj(a); }
__device__ inline double normD(double a) { return a * a; }
__device__ inline double normD(cd a) { return a.re * a.re + a.im * a.im; }
__device__ inline double recipD(double a) { return 1.0 / a; }
__device__ inline cd recipD(cd a) { return nq64::div(nq64::one(), a); }
__device__ inline bool isZeroD(double a) { return a == 0.0; }
__device__ inline bool isZeroD(cd a) { return a.re == 0.0 && a.im == 0.0; }
template <typename D> __device__ inline D oneD();
template <> __device__ inline double oneD<double>() { return 1.0; }
template <> __device__ inline cd oneD<cd>() { return nq64::one(); }
template <typename D> __device__ inline D zeroD();
template <> __device__ inline double zeroD<double>() { return 0.0; }
template <> __device__ inline cd zeroD<cd>() { return nq64::zero(); }

// FP32 <-> FP64 scalar of the same kind
__device__ inline double toD(float a) { return (double)a; }
__device__ inline cd toD(cf a) { return nq64::mk((double)a.re, (double)a.im); }
__device__ inline float toT(double a) { return (float)a; }
__device__ inline cf toT(cd a) { return mkf((float)a.re, (float)a.im); }

// lane shuffles
__device__ inline double shflxD(double v, int m) { return __shfl_xor(v, m); }
__device__ inline cd shflxD(cd v, int m) { return nq64::mk(__shfl_xor(v.re, m), __shfl_xor(v.im, m)); }

// Lane exchange x <- x[lane ^ MASK] without the LDS crossbar: DPP quad permutes
// (MASK 1, 2), row half-mirror (7) and row mirror (15) within 16-lane rows, and the
// gfx950 permlane swaps across rows (16) and halves (32).
template <int MASK>
__device__ inline uint32_t xor_lane(uint32_t x)
{
    if constexpr (MASK == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0xB1, 0xF, 0xF, false);
    } else if constexpr (MASK == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x4E, 0xF, 0xF, false);
    } else if constexpr (MASK == 7) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x141, 0xF, 0xF, false);
    } else if constexpr (MASK == 15) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, 0x140, 0xF, 0xF, false);
    } else if constexpr (MASK == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);  // rows [0,0,2,2] / [1,1,3,3]
        return (__lane_id() & 16) ? r[0] : r[1];
    } else {
        static_assert(MASK == 32, "unsupported lane mask");
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);  // halves [0,0] / [1,1]
        return (__lane_id() & 32) ? r[0] : r[1];
    }
}
template <int MASK>
__device__ inline double xor_lane(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t r = (uint64_t)xor_lane<MASK>((uint32_t)b) | ((uint64_t)xor_lane<MASK>((uint32_t)(b >> 32)) << 32);
    return __builtin_bit_cast(double, r);
}
template <int MASK>
__device__ inline cd xor_lane(cd v)
{
    return nq64::mk(xor_lane<MASK>(v.re), xor_lane<MASK>(v.im));
}

// Product / sum over the G lanes of a lane group (G a power of two <= 64) by a
// butterfly over the lane masks 1, 2, 7, 15, 16, 32 (a basis of the group's index
// space): lanes i and i ^ m combine the same two values at every step, so every lane
// of the group ends with the bitwise same result.  All lanes must be active.
template <int G, typename D>
__device__ inline D group_prod(D v)
{
    if constexpr (G >= 2) v = mulD(v, xor_lane<1>(v));
    if constexpr (G >= 4) v = mulD(v, xor_lane<2>(v));
    if constexpr (G >= 8) v = mulD(v, xor_lane<7>(v));
    if constexpr (G >= 16) v = mulD(v, xor_lane<15>(v));
    if constexpr (G >= 32) v = mulD(v, xor_lane<16>(v));
    if constexpr (G >= 64) v = mulD(v, xor_lane<32>(v));
    return v;
}
// FP32 exchange x[lane ^ MASK] within 8 lanes (MASK 1, 2, 7) written as a bare DPP
// move, which the compiler folds into the consuming multiply (one VALU op per step).
template <int MASK>
__device__ inline float xor_lane_f(float x)
{
    static_assert(MASK == 1 || MASK == 2 || MASK == 7, "unsupported lane mask");
    constexpr int ctrl = MASK == 1 ? 0xB1 : (MASK == 2 ? 0x4E : 0x141);
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), ctrl, 0xF, 0xF, true));
}
__device__ inline float mul_lane1(float v) { return v * xor_lane_f<1>(v); }
__device__ inline float mul_lane2(float v) { return v * xor_lane_f<2>(v); }
__device__ inline float mul_lane7(float v) { return v * xor_lane_f<7>(v); }
__device__ inline cf cmul_lanes(cf v, float re, float im) { return mkf(v.re * re - v.im * im, v.re * im + v.im * re); }
__device__ inline cf mul_lane1(cf v) { return cmul_lanes(v, xor_lane_f<1>(v.re), xor_lane_f<1>(v.im)); }
__device__ inline cf mul_lane2(cf v) { return cmul_lanes(v, xor_lane_f<2>(v.re), xor_lane_f<2>(v.im)); }
__device__ inline cf mul_lane7(cf v) { return cmul_lanes(v, xor_lane_f<7>(v.re), xor_lane_f<7>(v.im)); }

// Product over the G lanes of a group of per-lane FP32 (T) factors: the steps within
// 8 lanes in T, the rest in the FP64 type D.  With at most 8 factors per lane the T
// part multiplies at most 64 factors of magnitude e^{+-2|w|}, far inside the FP32
// range for any weights the solver reaches; the rounding it adds is of the same size
// as that of the FP32 factors themselves.  Every lane ends with the same bits.
// x[lane ^ MASK] through the LDS crossbar (ds_bpermute): no VALU work besides the
// consumer, for kernels bound by VALU issue
template <int MASK>
__device__ inline double xor_lane_lds(double v)
{
    const int addr = (int)((__lane_id() ^ MASK) << 2);
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)(uint32_t)(b >> 32));
    return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
}
template <int MASK>
__device__ inline cd xor_lane_lds(cd v)
{
    return nq64::mk(xor_lane_lds<MASK>(v.re), xor_lane_lds<MASK>(v.im));
}

// LDSX: the FP64 exchanges through ds_bpermute (fewer VALU instructions, LDS latency
// on the chain) instead of DPP / permlane moves
template <int G, typename T, typename D, bool LDSX = false>
__device__ inline D group_prod_tf(T v)
{
    if constexpr (G >= 2) v = mul_lane1(v);
    if constexpr (G >= 4) v = mul_lane2(v);
    if constexpr (G >= 8) v = mul_lane7(v);
    D d = toD(v);
    if constexpr (LDSX) {
        if constexpr (G >= 16) d = mulD(d, xor_lane_lds<15>(d));
        if constexpr (G >= 32) d = mulD(d, xor_lane_lds<16>(d));
        if constexpr (G >= 64) d = mulD(d, xor_lane_lds<32>(d));
    } else {
        if constexpr (G >= 16) d = mulD(d, xor_lane<15>(d));
        if constexpr (G >= 32) d = mulD(d, xor_lane<16>(d));
        if constexpr (G >= 64) d = mulD(d, xor_lane<32>(d));
    }
    return d;
}

template <int G, typename D>
__device__ inline D group_sum(D v)
{
    if constexpr (G >= 2) v = addD(v, xor_lane<1>(v));
    if constexpr (G >= 4) v = addD(v, xor_lane<2>(v));
    if constexpr (G >= 8) v = addD(v, xor_lane<7>(v));
    if constexpr (G >= 16) v = addD(v, xor_lane<15>(v));
    if constexpr (G >= 32) v = addD(v, xor_lane<16>(v));
    if constexpr (G >= 64) v = addD(v, xor_lane<32>(v));
    return v;
}

// ---------------------------------------------------------------- K x K in FP64
template <int K, typename D>
__device__ inline D det_m(const D *m)
{
    if constexpr (K == 1) {
        return m[0];
    } else if constexpr (K == 2) {
        return subD(mulD(m[0], m[3]), mulD(m[2], m[1]));
    } else if constexpr (K == 3) {
        D t0 = mulD(m[0], subD(mulD(m[4], m[8]), mulD(m[7], m[5])));
        D t1 = mulD(m[3], subD(mulD(m[1], m[8]), mulD(m[7], m[2])));
        D t2 = mulD(m[6], subD(mulD(m[1], m[5]), mulD(m[4], m[2])));
        return addD(subD(t0, t1), t2);
    } else {
        D acc = zeroD<D>();
#pragma unroll
        for (int r = 0; r < 4; r++) {
            D mn[9];
            int o = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (i == r) continue;
#pragma unroll
                for (int j = 1; j < 4; j++) mn[o++] = m[i * 4 + j];
            }
            D t = mulD(m[r * 4], det_m<3, D>(mn));
            acc = (r & 1) ? subD(acc, t) : addD(acc, t);
        }
        return acc;
    }
}

// inverse by adjugate with one reciprocal of det; singular -> zeros
template <int K, typename D>
__device__ inline void inv_m(const D *m, D *out)
{
    D d = det_m<K, D>(m);
    if (isZeroD(d)) {
#pragma unroll
        for (int i = 0; i < K * K; i++) out[i] = zeroD<D>();
        return;
    }
    D rd = recipD(d);
    if constexpr (K == 1) {
        out[0] = rd;
    } else {
#pragma unroll
        for (int i = 0; i < K; i++)
#pragma unroll
            for (int j = 0; j < K; j++) {
                // cofactor C_ji
                D mn[(K - 1) * (K - 1)];
                int o = 0;
#pragma unroll
                for (int r = 0; r < K; r++) {
                    if (r == j) continue;
#pragma unroll
                    for (int c = 0; c < K; c++) {
                        if (c == i) continue;
                        mn[o++] = m[r * K + c];
                    }
                }
                D v = det_m<K - 1, D>(mn);
                if ((i + j) & 1) v = negD(v);
                out[i * K + j] = mulD(v, rd);
            }
    }
}

template <int K, typename D>
__device__ inline void matmul_m(const D *A, const D *B, D *C)
{
#pragma unroll
    for (int r = 0; r < K; r++)
#pragma unroll
        for (int c = 0; c < K; c++) {
            D acc = mulD(A[r * K], B[c]);
#pragma unroll
            for (int l = 1; l < K; l++) acc = addD(acc, mulD(A[r * K + l], B[l * K + c]));
            C[r * K + c] = acc;
        }
}

} // namespace nqm
