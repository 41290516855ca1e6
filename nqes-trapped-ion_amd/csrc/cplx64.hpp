// cplx64.hpp -- FP64 complex arithmetic and elementary functions for the FP64
// device path.
//
// Every function here uses only IEEE-754 +, -, *, / (plus fabs/copysign and bit
// manipulation), and the translation units that include it are compiled with
// -ffp-contract=off, so results are bit-identical on any IEEE host or device that
// evaluates the same operations in the same order.  The operation order is the
// solver's canonical one:
//   complex multiply  (ac - bd, ad + bc)
//   complex divide    Smith (1962), one reciprocal of the scaled denominator
//   exp               Cody-Waite reduction by ln2 (32-bit high part), degree-13
//                     Taylor polynomial, scaling by 2^k
//   sin, cos          Cody-Waite reduction by pi/2 (three 33-bit parts) of |y|,
//                     Taylor polynomials to r^17 / r^18 on |r| <= pi/4
//   cosh, sinh        Taylor to x^20 / x^19 for |x| < 1, (e^x +- e^-x)/2 above
//   complex exp/cosh/sinh from the real ones; tanh(z) =
//                     (sinh x cosh x + i sin y cos y) / (sinh^2 x + cos^2 y), |x| <= 20
// sin/cosh/sinh are evaluated on |argument| and the sign applied last, so cosh and
// cos are exactly even and sinh and sin exactly odd; the sampler relies on that to
// tabulate cosh(2w), sinh(2w) once per SR step and flip signs per proposal.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nq64 {

struct cd {
    double re, im;
};

__host__ __device__ inline cd mk(double re, double im)
{
    cd z;
    z.re = re;
    z.im = im;
    return z;
}
__host__ __device__ inline cd add(cd a, cd b) { return mk(a.re + b.re, a.im + b.im); }
__host__ __device__ inline cd sub(cd a, cd b) { return mk(a.re - b.re, a.im - b.im); }
__host__ __device__ inline cd neg(cd a) { return mk(-a.re, -a.im); }
__host__ __device__ inline cd mul(cd a, cd b)
{
    return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ inline cd scale(double s, cd a) { return mk(s * a.re, s * a.im); }
__host__ __device__ inline cd divr(cd a, double s) { return mk(a.re / s, a.im / s); }
__host__ __device__ inline cd conj(cd a) { return mk(a.re, -a.im); }
__host__ __device__ inline double norm(cd a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ inline cd one() { return mk(1.0, 0.0); }
__host__ __device__ inline cd zero() { return mk(0.0, 0.0); }

__host__ __device__ inline cd div(cd a, cd b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re;
        double t = 1.0 / (b.re + b.im * r);
        return mk((a.re + a.im * r) * t, (a.im - a.re * r) * t);
    }
    double r = b.re / b.im;
    double t = 1.0 / (b.re * r + b.im);
    return mk((a.re * r + a.im) * t, (a.im * r - a.re) * t);
}

// ---------------------------------------------------------------- real functions
#define NQ_LN2_HI 0x1.62e42ff000000p-1
#define NQ_LN2_LO (-0x1.718432a1b0e26p-35)
#define NQ_INV_LN2 0x1.71547652b82fep+0
#define NQ_PIO2_1 0x1.921fb54400000p+0
#define NQ_PIO2_2 0x1.0b4611a600000p-34
#define NQ_PIO2_3 0x1.3198a2e037073p-69
#define NQ_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define NQ_RND_MAGIC 6755399441055744.0

// 1/n!, n = 0..20 (rounded quotients of exact integers)
#define NQ_F(n) nq_invfact_##n
#define NQ_DEF(n, v) static constexpr double nq_invfact_##n = 1.0 / v;
NQ_DEF(2, 2.0)
NQ_DEF(3, 6.0)
NQ_DEF(4, 24.0)
NQ_DEF(5, 120.0)
NQ_DEF(6, 720.0)
NQ_DEF(7, 5040.0)
NQ_DEF(8, 40320.0)
NQ_DEF(9, 362880.0)
NQ_DEF(10, 3628800.0)
NQ_DEF(11, 39916800.0)
NQ_DEF(12, 479001600.0)
NQ_DEF(13, 6227020800.0)
NQ_DEF(14, 87178291200.0)
NQ_DEF(15, 1307674368000.0)
NQ_DEF(16, 20922789888000.0)
NQ_DEF(17, 355687428096000.0)
NQ_DEF(18, 6402373705728000.0)
NQ_DEF(19, 121645100408832000.0)
NQ_DEF(20, 2432902008176640000.0)
#undef NQ_DEF

__host__ __device__ inline double rint_magic(double x) { return (x + NQ_RND_MAGIC) - NQ_RND_MAGIC; }

__host__ __device__ inline double pow2i(int n)
{
    uint64_t u = (uint64_t)(n + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &u, sizeof(d));
    return d;
}

__host__ __device__ inline double exp_(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return __builtin_huge_val();
    if (x < -745.1332191019412) return 0.0;
    double kd = rint_magic(x * NQ_INV_LN2);
    int k = (int)kd;
    double r = (x - kd * NQ_LN2_HI) - kd * NQ_LN2_LO;
    double p = NQ_F(13);
    p = NQ_F(12) + r * p;
    p = NQ_F(11) + r * p;
    p = NQ_F(10) + r * p;
    p = NQ_F(9) + r * p;
    p = NQ_F(8) + r * p;
    p = NQ_F(7) + r * p;
    p = NQ_F(6) + r * p;
    p = NQ_F(5) + r * p;
    p = NQ_F(4) + r * p;
    p = NQ_F(3) + r * p;
    p = NQ_F(2) + r * p;
    double e = 1.0 + (r + r * (r * p));
    if (k > 1023) {
        e = e * pow2i(1023);
        k -= 1023;
    }
    if (k < -1022) {
        e = e * pow2i(-1022);
        k += 1022;
    }
    if (k < -1022) {
        e = e * pow2i(-1022);
        k += 1022;
    }
    return e * pow2i(k);
}

__host__ __device__ inline void sincos_(double y, double *s, double *c)
{
    double ay = fabs(y);
    double q = rint_magic(ay * NQ_TWO_OVER_PI);
    double r = ((ay - q * NQ_PIO2_1) - q * NQ_PIO2_2) - q * NQ_PIO2_3;
    double z = r * r;
    double ps = -NQ_F(17);
    ps = NQ_F(15) + z * ps;
    ps = -NQ_F(13) + z * ps;
    ps = NQ_F(11) + z * ps;
    ps = -NQ_F(9) + z * ps;
    ps = NQ_F(7) + z * ps;
    ps = -NQ_F(5) + z * ps;
    ps = NQ_F(3) + z * ps;
    double sr = r - r * (z * ps);
    double pc = NQ_F(18);
    pc = -NQ_F(16) + z * pc;
    pc = NQ_F(14) + z * pc;
    pc = -NQ_F(12) + z * pc;
    pc = NQ_F(10) + z * pc;
    pc = -NQ_F(8) + z * pc;
    pc = NQ_F(6) + z * pc;
    pc = -NQ_F(4) + z * pc;
    double cr = 1.0 - (0.5 * z + (z * z) * pc);
    int quad = (int)((int64_t)q & 3);
    double so, co;
    if (quad == 0) {
        so = sr;
        co = cr;
    } else if (quad == 1) {
        so = cr;
        co = -sr;
    } else if (quad == 2) {
        so = -sr;
        co = -cr;
    } else {
        so = -cr;
        co = sr;
    }
    *s = (y < 0) ? -so : so;
    *c = co;
}

__host__ __device__ inline void coshsinh_(double x, double *ch, double *sh)
{
    double ax = fabs(x);
    double c, s;
    if (ax < 1.0) {
        double z = ax * ax;
        double ps = NQ_F(19);
        ps = NQ_F(17) + z * ps;
        ps = NQ_F(15) + z * ps;
        ps = NQ_F(13) + z * ps;
        ps = NQ_F(11) + z * ps;
        ps = NQ_F(9) + z * ps;
        ps = NQ_F(7) + z * ps;
        ps = NQ_F(5) + z * ps;
        ps = NQ_F(3) + z * ps;
        s = ax + ax * (z * ps);
        double pc = NQ_F(20);
        pc = NQ_F(18) + z * pc;
        pc = NQ_F(16) + z * pc;
        pc = NQ_F(14) + z * pc;
        pc = NQ_F(12) + z * pc;
        pc = NQ_F(10) + z * pc;
        pc = NQ_F(8) + z * pc;
        pc = NQ_F(6) + z * pc;
        pc = NQ_F(4) + z * pc;
        pc = NQ_F(2) + z * pc;
        c = 1.0 + z * pc;
    } else if (ax <= 709.0) {
        double e = exp_(ax);
        double ie = 1.0 / e;
        c = 0.5 * (e + ie);
        s = 0.5 * (e - ie);
    } else {
        double t = exp_(0.5 * ax);
        c = (0.5 * t) * t;
        s = c;
    }
    *ch = c;
    *sh = (x < 0) ? -s : s;
}

__host__ __device__ inline cd cexp_(cd z)
{
    double e = exp_(z.re), s, c;
    sincos_(z.im, &s, &c);
    return mk(e * c, e * s);
}

__host__ __device__ inline cd ccosh_(cd z)
{
    double ch, sh, s, c;
    coshsinh_(z.re, &ch, &sh);
    sincos_(z.im, &s, &c);
    return mk(ch * c, sh * s);
}

__host__ __device__ inline cd csinh_(cd z)
{
    double ch, sh, s, c;
    coshsinh_(z.re, &ch, &sh);
    sincos_(z.im, &s, &c);
    return mk(sh * c, ch * s);
}

__host__ __device__ inline cd ctanh_(cd z)
{
    double sy, cy;
    sincos_(z.im, &sy, &cy);
    if (fabs(z.re) > 20.0) {
        double e = exp_(-2.0 * fabs(z.re));
        return mk(__builtin_copysign(1.0, z.re), 4.0 * (sy * cy) * e);
    }
    double ch, sh;
    coshsinh_(z.re, &ch, &sh);
    double d = sh * sh + cy * cy;
    return mk((sh * ch) / d, (sy * cy) / d);
}

#undef NQ_F

// ---------------------------------------------------------------- K x K algebra
// Determinant by cofactor expansion along column 0; inverse = adjugate / det.
__host__ __device__ inline cd det2(cd a, cd b, cd c, cd d) { return sub(mul(a, d), mul(c, b)); }

template <int LD>
__host__ __device__ inline cd det3v(const cd *m)
{
    cd t0 = mul(m[0 * LD + 0], det2(m[1 * LD + 1], m[1 * LD + 2], m[2 * LD + 1], m[2 * LD + 2]));
    cd t1 = mul(m[1 * LD + 0], det2(m[0 * LD + 1], m[0 * LD + 2], m[2 * LD + 1], m[2 * LD + 2]));
    cd t2 = mul(m[2 * LD + 0], det2(m[0 * LD + 1], m[0 * LD + 2], m[1 * LD + 1], m[1 * LD + 2]));
    return add(sub(t0, t1), t2);
}

template <int K>
__host__ __device__ inline void minor_of(const cd *m, int r, int c, cd *out)
{
    int o = 0;
#pragma unroll
    for (int i = 0; i < K; i++) {
        if (i == r) continue;
#pragma unroll
        for (int j = 0; j < K; j++) {
            if (j == c) continue;
            out[o++] = m[i * K + j];
        }
    }
}

template <int K>
__host__ __device__ inline cd detK(const cd *m)
{
    if constexpr (K == 1) {
        return m[0];
    } else if constexpr (K == 2) {
        return det2(m[0], m[1], m[2], m[3]);
    } else if constexpr (K == 3) {
        return det3v<3>(m);
    } else {
        static_assert(K == 4, "K <= 4");
        cd acc = zero();
        cd mn[9];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            minor_of<4>(m, r, 0, mn);
            cd t = mul(m[r * 4], det3v<3>(mn));
            acc = (r == 0) ? t : ((r & 1) ? sub(acc, t) : add(acc, t));
        }
        return acc;
    }
}

template <int K>
__host__ __device__ inline cd cofactor(const cd *m, int r, int c)
{
    cd v;
    if constexpr (K == 1) {
        return one();
    } else if constexpr (K == 2) {
        cd mn[1];
        minor_of<2>(m, r, c, mn);
        v = mn[0];
    } else if constexpr (K == 3) {
        cd mn[4];
        minor_of<3>(m, r, c, mn);
        v = det2(mn[0], mn[1], mn[2], mn[3]);
    } else {
        cd mn[9];
        minor_of<4>(m, r, c, mn);
        v = det3v<3>(mn);
    }
    return ((r + c) & 1) ? neg(v) : v;
}

// inverse = adj / det; exactly singular -> zeros (the sampler then accepts the next
// proposal unconditionally and recomputes the inverse).
template <int K>
__host__ __device__ inline void invK(const cd *m, cd *out)
{
    cd d = detK<K>(m);
    if (d.re == 0.0 && d.im == 0.0) {
#pragma unroll
        for (int i = 0; i < K * K; i++) out[i] = zero();
        return;
    }
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < K; j++) out[i * K + j] = div(cofactor<K>(m, j, i), d);
}

// C = A B, inner index ascending from the first product
template <int K>
__host__ __device__ inline void matmulK(const cd *A, const cd *B, cd *C)
{
#pragma unroll
    for (int r = 0; r < K; r++)
#pragma unroll
        for (int c = 0; c < K; c++) {
            cd acc = mul(A[r * K], B[c]);
#pragma unroll
            for (int l = 1; l < K; l++) acc = add(acc, mul(A[r * K + l], B[l * K + c]));
            C[r * K + c] = acc;
        }
}

// ---------------------------------------------------------------- pcg64
// setseq_xsl_rr_128_64 with the default increment, state as two 64-bit halves.
struct pcg {
    uint64_t lo, hi;
};

#define NQ_PCG_MULT_HI 2549297995355413924ULL
#define NQ_PCG_MULT_LO 4865540595714422341ULL
#define NQ_PCG_INC_HI 6364136223846793005ULL
#define NQ_PCG_INC_LO 1442695040888963407ULL

__host__ __device__ inline uint64_t umulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// s <- s * MULT + INC (mod 2^128)
__host__ __device__ inline void pcg_bump(pcg &s)
{
    uint64_t lo = s.lo * NQ_PCG_MULT_LO;
    uint64_t hi = umulhi64(s.lo, NQ_PCG_MULT_LO) + s.lo * NQ_PCG_MULT_HI + s.hi * NQ_PCG_MULT_LO;
    uint64_t nlo = lo + NQ_PCG_INC_LO;
    hi += NQ_PCG_INC_HI + (nlo < lo ? 1u : 0u);
    s.lo = nlo;
    s.hi = hi;
}

// pcg64(seed): state = bump(seed + INC)
__host__ __device__ inline pcg pcg_seed(uint64_t seed)
{
    pcg s;
    s.lo = seed + NQ_PCG_INC_LO;
    s.hi = NQ_PCG_INC_HI + (s.lo < seed ? 1u : 0u);
    pcg_bump(s);
    return s;
}

// advance, then XSL-RR of the new state
__host__ __device__ inline uint64_t pcg_next(pcg &s)
{
    pcg_bump(s);
    unsigned rot = (unsigned)(s.hi >> 58);
    uint64_t x = s.hi ^ s.lo;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// XSL-RR output of a state
__host__ __device__ inline uint64_t pcg_out(const pcg &s)
{
    unsigned rot = (unsigned)(s.hi >> 58);
    uint64_t x = s.hi ^ s.lo;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// jump-ahead: the affine map s -> a s + c (mod 2^128) applied d times is
// s -> a_d s + c_d; (a, c) holds (a_d, c_d).  pcg_affine applies it.
__host__ __device__ inline pcg pcg_affine(const pcg &s, const pcg &a, const pcg &c)
{
    uint64_t lo = s.lo * a.lo;
    uint64_t hi = umulhi64(s.lo, a.lo) + s.lo * a.hi + s.hi * a.lo;
    uint64_t nlo = lo + c.lo;
    hi += c.hi + (nlo < lo ? 1u : 0u);
    pcg r;
    r.lo = nlo;
    r.hi = hi;
    return r;
}

// (a_d, c_d) of d >= 1 steps of the pcg64 state transition
__host__ __device__ inline void pcg_jump(unsigned d, pcg &a, pcg &c)
{
    const pcg A{NQ_PCG_MULT_LO, NQ_PCG_MULT_HI}, C{NQ_PCG_INC_LO, NQ_PCG_INC_HI};
    a = A;
    c = C;
    for (unsigned k = 1; k < d; k++) {  // f o f^k: (A a_k, A c_k + C)
        const pcg zero{0, 0};
        a = pcg_affine(a, A, zero);
        c = pcg_affine(c, A, C);
    }
}

// generate_canonical<double, 53> over a 64-bit engine
__host__ __device__ inline double canonical(uint64_t u)
{
    double r = (double)u * 0x1p-64;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
}

} // namespace nq64
