/*
 * nqes_oracle.c -- CPU oracle for the multi-RBM variational Monte Carlo SR step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the smoke
 * check and the CPU-baseline leg of bench.py load it (as oracle/liboracle.so)
 * to compare the HIP path against.  Nothing in the product links or calls it.
 *
 * It is a clean-room restatement, in plain C with FP64 arithmetic and OpenMP
 * over chains, of the reference's hot path (reference file code/rbm.cc):
 *   - pcg64 seeding/stepping                    code/pcg_random.h:161-164, 484-488, 431-437, 1060-1086
 *   - generate_canonical / bernoulli / uniform  libstdc++ <random> (see orc_canonical)
 *   - init_multi_spin                           code/rbm.cc:99-119
 *   - theta / state_fcn                         code/rbm.cc:135-173
 *   - flip_state_fcn                            code/rbm.cc:176-197
 *   - deriv                                     code/rbm.cc:199-239
 *   - monte_carlo_step                          code/rbm.cc:530-552
 *   - monte_carlo (SR step, matrix-free)        code/rbm.cc:564-720
 *   - trapped_ion_model                         code/rbm.cc:1012-1031
 *   - ising_like_hamiltonian                    code/rbm.cc:1042-1064
 *   - hs_like_hamiltonian (models hs, xxz)      code/rbm.cc:1066-1119
 *   - minresqlp / sym_ortho                     code/minresqlp.h:85-390
 *
 * Canonical floating-point order.  The reference's exact bits depend on glibc's
 * complex libm, Eigen's kernels and the compiler's FMA contraction, none of which
 * can be reproduced here.  The oracle therefore fixes one operation order and the
 * HIP FP64 path follows the same order, so the two agree bit for bit:
 *   - elementary functions are the portable implementations below (Cody-Waite
 *     reduction + Taylor polynomials), built only from IEEE +,-,*,/;
 *   - complex multiply is (ac-bd, ad+bc); complex divide is Smith's (1962)
 *     algorithm with one reciprocal;
 *   - the product of the M per-hidden-unit factors (rbm.cc:159-168, 189-194) is
 *     taken as: lane partial g = prod_{j = g mod 64} d_j in ascending j, then a
 *     pairwise tree over the 64 partials with strides 32,16,...,1 (missing
 *     partials are exactly 1).  The reference multiplies left to right; this is a
 *     reassociation of the same factors;
 *   - K x K determinants and inverses use cofactor expansion along column 0
 *     (closed form, K <= 4);
 *   - every other sum runs in ascending index order.
 * Build with -ffp-contract=off (no FMA contraction) and without -ffast-math.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_KMAX 4
#define ORC_NMAX 1024

/* ------------------------------------------------------------------------- */
/* complex arithmetic (canonical)                                            */
/* ------------------------------------------------------------------------- */
typedef struct { double re, im; } cplx;

static inline cplx cmk(double re, double im) { cplx z; z.re = re; z.im = im; return z; }
static inline cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cplx cneg(cplx a) { return cmk(-a.re, -a.im); }
static inline cplx cmul(cplx a, cplx b)
{
    return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cplx cscale(double s, cplx a) { return cmk(s * a.re, s * a.im); }
static inline cplx cdivr(cplx a, double s) { return cmk(a.re / s, a.im / s); }
static inline cplx cconj(cplx a) { return cmk(a.re, -a.im); }
static inline double cnorm(cplx a) { return a.re * a.re + a.im * a.im; }

/* Smith's algorithm with a single reciprocal of the scaled denominator. */
static inline cplx cdiv(cplx a, cplx b)
{
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re;
        double t = 1.0 / (b.re + b.im * r);
        return cmk((a.re + a.im * r) * t, (a.im - a.re * r) * t);
    } else {
        double r = b.re / b.im;
        double t = 1.0 / (b.re * r + b.im);
        return cmk((a.re * r + a.im) * t, (a.im * r - a.re) * t);
    }
}

/* ------------------------------------------------------------------------- */
/* elementary functions: only IEEE +,-,*,/ so that the device copy matches   */
/* ------------------------------------------------------------------------- */
static const double LN2_HI = 0x1.62e42ff000000p-1;     /* ln2 to 32 bits      */
static const double LN2_LO = -0x1.718432a1b0e26p-35;   /* ln2 - LN2_HI        */
static const double INV_LN2 = 0x1.71547652b82fep+0;
static const double PIO2_1 = 0x1.921fb54400000p+0;     /* pi/2 to 33 bits     */
static const double PIO2_2 = 0x1.0b4611a600000p-34;    /* next 33 bits        */
static const double PIO2_3 = 0x1.3198a2e037073p-69;    /* remainder           */
static const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
static const double RND_MAGIC = 6755399441055744.0;    /* 1.5 * 2^52          */

/* 1/n! for n = 0..20, rounded to double */
static const double INVFACT[21] = {
    1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0,
    1.0 / 5040.0, 1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0,
    1.0 / 39916800.0, 1.0 / 479001600.0, 1.0 / 6227020800.0,
    1.0 / 87178291200.0, 1.0 / 1307674368000.0, 1.0 / 20922789888000.0,
    1.0 / 355687428096000.0, 1.0 / 6402373705728000.0,
    1.0 / 121645100408832000.0, 1.0 / 2432902008176640000.0
};

/* round-to-nearest-even integer of x, for |x| < 2^51 */
static inline double rint_magic(double x) { return (x + RND_MAGIC) - RND_MAGIC; }

/* 2^n as a double for -1022 <= n <= 1023 */
static inline double pow2i(int n)
{
    union { uint64_t u; double d; } v;
    v.u = (uint64_t)(n + 1023) << 52;
    return v.d;
}

double orc_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    double kd = rint_magic(x * INV_LN2);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    /* Taylor polynomial of degree 13 on |r| <= ln2/2 */
    double p = INVFACT[13];
    for (int n = 12; n >= 2; n--) p = INVFACT[n] + r * p;
    double e = 1.0 + (r + r * (r * p));
    /* scale by 2^k in at most two steps (k may be outside the normal range) */
    if (k > 1023) { e = e * pow2i(1023); k -= 1023; }
    if (k < -1022) { e = e * pow2i(-1022); k += 1022; }
    if (k < -1022) { e = e * pow2i(-1022); k += 1022; }
    return e * pow2i(k);
}

/* sin and cos of r, |r| <= pi/4 */
static inline void sincos_poly(double r, double *s, double *c)
{
    double z = r * r;
    double ps = -INVFACT[17];
    ps = INVFACT[15] + z * ps;
    ps = -INVFACT[13] + z * ps;
    ps = INVFACT[11] + z * ps;
    ps = -INVFACT[9] + z * ps;
    ps = INVFACT[7] + z * ps;
    ps = -INVFACT[5] + z * ps;
    ps = INVFACT[3] + z * ps;       /* = 1/6 - ...               */
    *s = r - r * (z * ps);          /* r - r^3/6 + r^5/120 - ... */
    double pc = INVFACT[18];
    pc = -INVFACT[16] + z * pc;
    pc = INVFACT[14] + z * pc;
    pc = -INVFACT[12] + z * pc;
    pc = INVFACT[10] + z * pc;
    pc = -INVFACT[8] + z * pc;
    pc = INVFACT[6] + z * pc;
    pc = -INVFACT[4] + z * pc;      /* = -1/24 + z/720 - ...     */
    *c = 1.0 - (0.5 * z + (z * z) * pc);
}

/* sin/cos built on |y| so that sin is exactly odd and cos exactly even. */
void orc_sincos(double y, double *s, double *c)
{
    double ay = fabs(y);
    double q = rint_magic(ay * TWO_OVER_PI);
    double r = ((ay - q * PIO2_1) - q * PIO2_2) - q * PIO2_3;
    double sr, cr;
    sincos_poly(r, &sr, &cr);
    int quad = (int)((int64_t)q & 3);
    double so, co;
    switch (quad) {
    case 0: so = sr; co = cr; break;
    case 1: so = cr; co = -sr; break;
    case 2: so = -sr; co = -cr; break;
    default: so = -cr; co = sr; break;
    }
    *s = (y < 0) ? -so : so;
    *c = co;
}

/* cosh and sinh together; exactly even/odd by construction. */
void orc_coshsinh(double x, double *ch, double *sh)
{
    double ax = fabs(x);
    double c, s;
    if (ax < 1.0) {
        double z = ax * ax;
        double ps = INVFACT[19];
        for (int n = 17; n >= 3; n -= 2) ps = INVFACT[n] + z * ps;
        s = ax + ax * (z * ps);
        double pc = INVFACT[20];
        for (int n = 18; n >= 2; n -= 2) pc = INVFACT[n] + z * pc;
        c = 1.0 + z * pc;
    } else if (ax <= 709.0) {
        double e = orc_exp(ax);
        double ie = 1.0 / e;
        c = 0.5 * (e + ie);
        s = 0.5 * (e - ie);
    } else {
        double t = orc_exp(0.5 * ax);
        c = (0.5 * t) * t;
        s = c;
    }
    *ch = c;
    *sh = (x < 0) ? -s : s;
}

static inline cplx c_exp(cplx z)
{
    double e = orc_exp(z.re), s, c;
    orc_sincos(z.im, &s, &c);
    return cmk(e * c, e * s);
}

static inline cplx c_cosh(cplx z)
{
    double ch, sh, s, c;
    orc_coshsinh(z.re, &ch, &sh);
    orc_sincos(z.im, &s, &c);
    return cmk(ch * c, sh * s);
}

static inline cplx c_sinh(cplx z)
{
    double ch, sh, s, c;
    orc_coshsinh(z.re, &ch, &sh);
    orc_sincos(z.im, &s, &c);
    return cmk(sh * c, ch * s);
}

/* tanh(x+iy) = (sinh x cosh x + i sin y cos y) / (sinh^2 x + cos^2 y): the
 * denominator equals (cosh 2x + cos 2y)/2 without its cancellation near the poles */
static inline cplx c_tanh(cplx z)
{
    double sy, cy;
    orc_sincos(z.im, &sy, &cy);
    if (fabs(z.re) > 20.0) {
        double e = orc_exp(-2.0 * fabs(z.re));
        return cmk(copysign(1.0, z.re), 4.0 * (sy * cy) * e);
    }
    double ch, sh;
    orc_coshsinh(z.re, &ch, &sh);
    double d = sh * sh + cy * cy;
    return cmk((sh * ch) / d, (sy * cy) / d);
}

/* exported scalar wrappers (tests pin them against numpy) */
void orc_sincos_v(double y, double *out) { orc_sincos(y, out, out + 1); }
void orc_coshsinh_v(double x, double *out) { orc_coshsinh(x, out, out + 1); }
void orc_cexp(const double *z, double *o) { cplx r = c_exp(cmk(z[0], z[1])); o[0] = r.re; o[1] = r.im; }
void orc_ccosh(const double *z, double *o) { cplx r = c_cosh(cmk(z[0], z[1])); o[0] = r.re; o[1] = r.im; }
void orc_csinh(const double *z, double *o) { cplx r = c_sinh(cmk(z[0], z[1])); o[0] = r.re; o[1] = r.im; }
void orc_ctanh(const double *z, double *o) { cplx r = c_tanh(cmk(z[0], z[1])); o[0] = r.re; o[1] = r.im; }
void orc_cdiv(const double *a, const double *b, double *o)
{
    cplx r = cdiv(cmk(a[0], a[1]), cmk(b[0], b[1]));
    o[0] = r.re; o[1] = r.im;
}

/* canonical product of M factors (see header comment) */
static cplx tree_prod(const cplx *d, int M)
{
    cplx part[64];
    for (int g = 0; g < 64; g++) {
        if (g < M) {
            cplx p = d[g];
            for (int j = g + 64; j < M; j += 64) p = cmul(p, d[j]);
            part[g] = p;
        } else {
            part[g] = cmk(1.0, 0.0);
        }
    }
    for (int st = 32; st >= 1; st >>= 1)
        for (int g = 0; g < st; g++) part[g] = cmul(part[g], part[g + st]);
    return part[0];
}

void orc_tree_prod(const double *d, int M, double *out)
{
    cplx r = tree_prod((const cplx *)d, M);
    out[0] = r.re; out[1] = r.im;
}

/* ------------------------------------------------------------------------- */
/* PCG64 (setseq_xsl_rr_128_64) and libstdc++ distributions                  */
/* ------------------------------------------------------------------------- */
typedef unsigned __int128 u128;
static const u128 PCG_MULT = ((u128)2549297995355413924ULL << 64) | 4865540595714422341ULL;
static const u128 PCG_INC = ((u128)6364136223846793005ULL << 64) | 1442695040888963407ULL;

/* pcg64(seed): state = bump(seed + inc), pcg_random.h:484-488 */
static inline u128 pcg_seed(uint64_t seed) { return ((u128)seed + PCG_INC) * PCG_MULT + PCG_INC; }

/* operator(): advance, then XSL-RR output of the new state (128-bit state
 * uses output_previous = false), pcg_random.h:431-437, 1060-1086 */
static inline uint64_t pcg_next(u128 *st)
{
    *st = *st * PCG_MULT + PCG_INC;
    u128 s = *st;
    unsigned rot = (unsigned)(s >> 122) & 63u;
    uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

/* std::generate_canonical<double,53> over a 64-bit engine: one draw,
 * U = double(u) / 2^64, clamped to nextafter(1, 0) (libstdc++ random.tcc). */
static inline double canonical(uint64_t u)
{
    double r = (double)u * 0x1p-64;
    if (r >= 1.0) r = 0x1.fffffffffffffp-1;
    return r;
}

void orc_pcg64_init(uint64_t seed, uint64_t *st)
{
    u128 s = pcg_seed(seed);
    st[0] = (uint64_t)s; st[1] = (uint64_t)(s >> 64);
}

uint64_t orc_pcg64_next(uint64_t *st)
{
    u128 s = ((u128)st[1] << 64) | st[0];
    uint64_t r = pcg_next(&s);
    st[0] = (uint64_t)s; st[1] = (uint64_t)(s >> 64);
    return r;
}

double orc_canonical(uint64_t u) { return canonical(u); }

/* ------------------------------------------------------------------------- */
/* K x K determinant / inverse (cofactor expansion along column 0)           */
/* ------------------------------------------------------------------------- */
static cplx det2(cplx a, cplx b, cplx c, cplx d)  /* | a b ; c d | */
{
    return csub(cmul(a, d), cmul(c, b));
}

static cplx det3v(const cplx *m, int ld)
{
#define E(r, c) m[(r) * ld + (c)]
    cplx t0 = cmul(E(0, 0), det2(E(1, 1), E(1, 2), E(2, 1), E(2, 2)));
    cplx t1 = cmul(E(1, 0), det2(E(0, 1), E(0, 2), E(2, 1), E(2, 2)));
    cplx t2 = cmul(E(2, 0), det2(E(0, 1), E(0, 2), E(1, 1), E(1, 2)));
#undef E
    return cadd(csub(t0, t1), t2);
}

/* minor of a K x K matrix with row r and column c removed, stored (K-1)x(K-1) */
static void minor_of(const cplx *m, int K, int r, int c, cplx *out)
{
    int o = 0;
    for (int i = 0; i < K; i++) {
        if (i == r) continue;
        for (int j = 0; j < K; j++) {
            if (j == c) continue;
            out[o++] = m[i * K + j];
        }
    }
}

static cplx detK(const cplx *m, int K)
{
    switch (K) {
    case 1: return m[0];
    case 2: return det2(m[0], m[1], m[2], m[3]);
    case 3: return det3v(m, 3);
    default: {
        cplx acc = cmk(0.0, 0.0);
        cplx mn[9];
        for (int r = 0; r < 4; r++) {
            minor_of(m, 4, r, 0, mn);
            cplx t = cmul(m[r * 4], det3v(mn, 3));
            acc = (r == 0) ? t : ((r & 1) ? csub(acc, t) : cadd(acc, t));
        }
        return acc;
    }
    }
}

/* cofactor C_rc = (-1)^(r+c) det(minor_rc) */
static cplx cofactor(const cplx *m, int K, int r, int c)
{
    cplx v;
    if (K == 1) return cmk(1.0, 0.0);
    if (K == 2) {
        cplx mn[1];
        minor_of(m, 2, r, c, mn);
        v = mn[0];
    } else if (K == 3) {
        cplx mn[4];
        minor_of(m, 3, r, c, mn);
        v = det2(mn[0], mn[1], mn[2], mn[3]);
    } else {
        cplx mn[9];
        minor_of(m, 4, r, c, mn);
        v = det3v(mn, 3);
    }
    return ((r + c) & 1) ? cneg(v) : v;
}

/* inverse = adj / det; exactly singular -> zeros (returns 0).  The sampler only
 * meets an exactly singular Phi at chain start when two replicas coincide; the
 * next proposal is then accepted unconditionally and the inverse recomputed. */
static int invK(const cplx *m, int K, cplx *out)
{
    cplx d = detK(m, K);
    if (d.re == 0.0 && d.im == 0.0) {
        for (int i = 0; i < K * K; i++) out[i] = cmk(0.0, 0.0);
        return 0;
    }
    for (int i = 0; i < K; i++)
        for (int j = 0; j < K; j++)
            out[i * K + j] = cdiv(cofactor(m, K, j, i), d);
    return 1;
}

void orc_det(int K, const double *m, double *out)
{
    cplx d = detK((const cplx *)m, K);
    out[0] = d.re; out[1] = d.im;
}

int orc_inv(int K, const double *m, double *out) { return invK((const cplx *)m, K, (cplx *)out); }

/* C = A * B, K x K, inner index ascending starting from the first product */
static void matmulK(const cplx *A, const cplx *B, int K, cplx *C)
{
    for (int r = 0; r < K; r++)
        for (int c = 0; c < K; c++) {
            cplx acc = cmul(A[r * K], B[c]);
            for (int l = 1; l < K; l++) acc = cadd(acc, cmul(A[r * K + l], B[l * K + c]));
            C[r * K + c] = acc;
        }
}

/* ------------------------------------------------------------------------- */
/* problem description and spins                                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int N, M, K, NW;    /* NW = 32-bit words per replica */
    long L, P;          /* L = N + M + N*M, P = K*L       */
    double h;
    const double *J;    /* N x N row-major                */
    int model;          /* 0 ising-like, 1 hs-like delta = 1, 2 hs-like delta = -1 */
} prob_t;

static prob_t mkprob(int N, int M, int K, double h, const double *J)
{
    prob_t p;
    p.N = N; p.M = M; p.K = K; p.NW = (N + 31) / 32;
    p.L = (long)N + M + (long)N * M; p.P = (long)K * p.L;
    p.h = h; p.J = J;
    p.model = 0;
    return p;
}

static inline int sbit(const uint32_t *s, int i) { return (int)((s[i >> 5] >> (i & 31)) & 1u); }
static inline void sflip(uint32_t *s, int i) { s[i >> 5] ^= 1u << (i & 31); }

/* weights: RBM kk occupies [kk*L, (kk+1)*L): a[N], b[M], w[N][M] */
static inline cplx Wa(const prob_t *p, const cplx *W, int kk, int i) { return W[kk * p->L + i]; }
static inline cplx Wb(const prob_t *p, const cplx *W, int kk, int j) { return W[kk * p->L + p->N + j]; }
static inline cplx Ww(const prob_t *p, const cplx *W, int kk, int i, int j)
{
    return W[kk * p->L + p->N + p->M + (long)i * p->M + j];
}

/* init_multi_spin: K*N bernoulli(0.5) draws, replica-major, bit i ascending
 * (rbm.cc:99-119); bernoulli(p) is U < p (libstdc++ random.h). */
static void init_spins(const prob_t *p, u128 *rng, uint32_t *s)
{
    memset(s, 0, sizeof(uint32_t) * p->K * p->NW);
    for (int k = 0; k < p->K; k++)
        for (int i = 0; i < p->N; i++)
            if (canonical(pcg_next(rng)) < 0.5) s[k * p->NW + (i >> 5)] |= 1u << (i & 31);
}

/* chain start spins of chain c: pcg64(seed + c), K*N draws (tests) */
void orc_init_spins(int N, int K, uint64_t seed, int c, uint32_t *out)
{
    prob_t p = mkprob(N, 1, K, 0.0, NULL);
    u128 rng = pcg_seed(seed + (uint64_t)c);
    init_spins(&p, &rng, out);
}

/* ------------------------------------------------------------------------- */
/* amplitudes                                                                */
/* ------------------------------------------------------------------------- */
/* tanh layout: tanh[(k*K + kk)*M + j] = tanh(theta_j of RBM kk at replica k) */

/* state_fcn (rbm.cc:148-173) */
static void state_fcn(const prob_t *p, const cplx *W, const uint32_t *s, cplx *phi, cplx *tanh_)
{
    int N = p->N, M = p->M, K = p->K;
    cplx *ch = (cplx *)malloc(sizeof(cplx) * M);
    for (int k = 0; k < K; k++) {
        const uint32_t *sk = s + k * p->NW;
        for (int kk = 0; kk < K; kk++) {
            cplx A = cmk(0.0, 0.0);
            for (int i = 0; i < N; i++) A = cadd(A, sbit(sk, i) ? cneg(Wa(p, W, kk, i)) : Wa(p, W, kk, i));
            for (int j = 0; j < M; j++) {
                cplx th = Wb(p, W, kk, j);
                for (int i = 0; i < N; i++) {
                    cplx w = Ww(p, W, kk, i, j);
                    th = cadd(th, sbit(sk, i) ? cneg(w) : w);
                }
                ch[j] = c_cosh(th);
                if (tanh_) tanh_[(k * K + kk) * M + j] = c_tanh(th);
            }
            phi[k * K + kk] = cmul(c_exp(A), tree_prod(ch, M));
        }
    }
    free(ch);
}

/* flip_state_fcn for replica k, spin i (rbm.cc:176-197): updates row k of phi
 * and the tanh columns (k, kk) in place.  scratch: M complex. */
static void flip_row(const prob_t *p, const cplx *W, int k, int i, const uint32_t *s,
                     cplx *phi, cplx *tanh_, cplx *scratch)
{
    int M = p->M, K = p->K;
    int si = sbit(s + k * p->NW, i);
    for (int kk = 0; kk < K; kk++) {
        cplx a2 = cscale(2.0, Wa(p, W, kk, i));
        cplx e = c_exp(si ? a2 : cneg(a2));
        cplx *t = tanh_ + (k * K + kk) * M;
        for (int j = 0; j < M; j++) {
            cplx w2 = cscale(2.0, Ww(p, W, kk, i, j));
            cplx ws = si ? cneg(w2) : w2;
            cplx chv = c_cosh(ws), shv = c_sinh(ws);
            cplx th = t[j];
            cplx denom = csub(chv, cmul(th, shv));
            t[j] = cdiv(csub(cmul(th, chv), shv), denom);
            scratch[j] = denom;
        }
        phi[k * K + kk] = cmul(cmul(phi[k * K + kk], e), tree_prod(scratch, M));
    }
}

/* trapped_ion_model (rbm.cc:1012-1031): 2 * sum_{i>j} J_ij s_i s_j */
static double ezz(const prob_t *p, const uint32_t *sk)
{
    double res = 0.0;
    for (int i = 1; i < p->N; i++)
        for (int j = 0; j < i; j++) {
            double v = p->J[(long)i * p->N + j];
            res += (sbit(sk, i) ^ sbit(sk, j)) ? -v : v;
        }
    return res * 2.0;
}

/* ising_like_hamiltonian (rbm.cc:1042-1064):
 * H = Phi^+ (diag(E_zz(s^k)) Phi - h sum_i Phi^(i)), Phi^(i) = Phi with spin i
 * flipped in every replica. */
static void hloc_ising(const prob_t *p, const cplx *W, const uint32_t *s, const cplx *phi,
                       const cplx *phi_inv, const cplx *tanh_, cplx *H, cplx *scratch)
{
    int N = p->N, M = p->M, K = p->K;
    cplx res[ORC_KMAX * ORC_KMAX], sx[ORC_KMAX * ORC_KMAX];
    for (int k = 0; k < K; k++) {
        double e = ezz(p, s + k * p->NW);
        for (int kk = 0; kk < K; kk++) res[k * K + kk] = cscale(e, phi[k * K + kk]);
    }
    for (int q = 0; q < K * K; q++) sx[q] = cmk(0.0, 0.0);
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < K; k++) {
            int si = sbit(s + k * p->NW, i);
            for (int kk = 0; kk < K; kk++) {
                cplx a2 = cscale(2.0, Wa(p, W, kk, i));
                cplx e = c_exp(si ? a2 : cneg(a2));
                const cplx *t = tanh_ + (k * K + kk) * M;
                for (int j = 0; j < M; j++) {
                    cplx w2 = cscale(2.0, Ww(p, W, kk, i, j));
                    cplx ws = si ? cneg(w2) : w2;
                    scratch[j] = csub(c_cosh(ws), cmul(t[j], c_sinh(ws)));
                }
                cplx v = cmul(cmul(phi[k * K + kk], e), tree_prod(scratch, M));
                sx[k * K + kk] = cadd(sx[k * K + kk], v);
            }
        }
    }
    for (int q = 0; q < K * K; q++) res[q] = csub(res[q], cscale(p->h, sx[q]));
    matmulK(phi_inv, res, K, H);
}

/* hs_like_hamiltonian (rbm.cc:1092-1119), H = sum_{i<j} J_ij (sx sx + sy sy + delta sz sz):
 * res row k = sum over pairs i > j (i ascending, then j ascending) of
 *   J_ij (2 Phi^(ij)_k - delta Phi_k)  if s^k_i != s^k_j,   (J_ij delta) Phi_k  else,
 * H = Phi^+ res.  Phi^(ij)_k is row k of Phi with spin i, then spin j flipped in
 * replica k: two flip_state_fcn steps (rbm.cc:1106-1108), the second on the angles
 * left by the first.  Pairs with J_ij == 0 add exact zeros and are skipped.
 * scratch: 2 M complex. */
static void hloc_hs(const prob_t *p, const cplx *W, const uint32_t *s, const cplx *phi,
                    const cplx *phi_inv, const cplx *tanh_, cplx *H, cplx *scratch)
{
    int N = p->N, M = p->M, K = p->K;
    double delta = p->model == 1 ? 1.0 : -1.0;
    cplx res[ORC_KMAX * ORC_KMAX];
    cplx *den = scratch, *t1 = scratch + M;
    for (int q = 0; q < K * K; q++) res[q] = cmk(0.0, 0.0);
    for (int i = 1; i < N; i++)
        for (int j = 0; j < i; j++) {
            double pij = p->J[(long)i * N + j];
            if (pij == 0.0) continue;
            for (int k = 0; k < K; k++) {
                int si = sbit(s + k * p->NW, i), sj = sbit(s + k * p->NW, j);
                if (si == sj) {
                    for (int kk = 0; kk < K; kk++)
                        res[k * K + kk] = cadd(res[k * K + kk], cscale(pij * delta, phi[k * K + kk]));
                    continue;
                }
                for (int kk = 0; kk < K; kk++) {
                    const cplx *t = tanh_ + (k * K + kk) * M;
                    cplx a2 = cscale(2.0, Wa(p, W, kk, i));
                    cplx e = c_exp(si ? a2 : cneg(a2));
                    for (int q = 0; q < M; q++) {
                        cplx w2 = cscale(2.0, Ww(p, W, kk, i, q));
                        cplx ws = si ? cneg(w2) : w2;
                        cplx chv = c_cosh(ws), shv = c_sinh(ws);
                        den[q] = csub(chv, cmul(t[q], shv));
                        t1[q] = cdiv(csub(cmul(t[q], chv), shv), den[q]);
                    }
                    cplx v = cmul(cmul(phi[k * K + kk], e), tree_prod(den, M));
                    cplx b2 = cscale(2.0, Wa(p, W, kk, j));
                    cplx e2 = c_exp(sj ? b2 : cneg(b2));
                    for (int q = 0; q < M; q++) {
                        cplx w2 = cscale(2.0, Ww(p, W, kk, j, q));
                        cplx ws = sj ? cneg(w2) : w2;
                        den[q] = csub(c_cosh(ws), cmul(t1[q], c_sinh(ws)));
                    }
                    v = cmul(cmul(v, e2), tree_prod(den, M));
                    cplx d = csub(cscale(2.0, v), cscale(delta, phi[k * K + kk]));
                    res[k * K + kk] = cadd(res[k * K + kk], cscale(pij, d));
                }
            }
        }
    matmulK(phi_inv, res, K, H);
}

/* the model's local energy; scratch: 2 M complex */
static void hloc(const prob_t *p, const cplx *W, const uint32_t *s, const cplx *phi,
                 const cplx *phi_inv, const cplx *tanh_, cplx *H, cplx *scratch)
{
    if (p->model) hloc_hs(p, W, s, phi, phi_inv, tanh_, H, scratch);
    else hloc_ising(p, W, s, phi, phi_inv, tanh_, H, scratch);
}

/* deriv (rbm.cc:199-239): O_p = det(Phi^+ derPhi), derPhi = Phi with column k
 * multiplied row-wise by d_kk (a: sigma_i, b: tanh theta_j, w: both). */
static inline cplx deriv_entry(const cplx *phi, const cplx *phi_inv, int K, int k,
                               const cplx *colk)
{
    cplx der[ORC_KMAX * ORC_KMAX], X[ORC_KMAX * ORC_KMAX];
    memcpy(der, phi, sizeof(cplx) * K * K);
    for (int kk = 0; kk < K; kk++) der[kk * K + k] = colk[kk];
    matmulK(phi_inv, der, K, X);
    return detK(X, K);
}

static void deriv(const prob_t *p, const uint32_t *s, const cplx *phi, const cplx *phi_inv,
                  const cplx *tanh_, cplx *O)
{
    int N = p->N, M = p->M, K = p->K;
    long L = p->L;
    cplx col[ORC_KMAX];
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < N; i++) {
            for (int kk = 0; kk < K; kk++) {
                cplx v = phi[kk * K + k];
                col[kk] = sbit(s + kk * p->NW, i) ? cneg(v) : v;
            }
            O[k * L + i] = deriv_entry(phi, phi_inv, K, k, col);
        }
        for (int j = 0; j < M; j++) {
            for (int kk = 0; kk < K; kk++)
                col[kk] = cmul(phi[kk * K + k], tanh_[(kk * K + k) * M + j]);
            O[k * L + N + j] = deriv_entry(phi, phi_inv, K, k, col);
        }
        for (int i = 0; i < N; i++)
            for (int j = 0; j < M; j++) {
                for (int kk = 0; kk < K; kk++) {
                    cplx v = cmul(phi[kk * K + k], tanh_[(kk * K + k) * M + j]);
                    col[kk] = sbit(s + kk * p->NW, i) ? cneg(v) : v;
                }
                O[k * L + N + M + (long)i * M + j] = deriv_entry(phi, phi_inv, K, k, col);
            }
    }
}

/* exported per-sample pieces */
void orc_state_fcn(int N, int M, int K, const double *W, const uint32_t *s, double *phi, double *tanh_)
{
    prob_t p = mkprob(N, M, K, 0.0, NULL);
    state_fcn(&p, (const cplx *)W, s, (cplx *)phi, (cplx *)tanh_);
}

void orc_flip(int N, int M, int K, const double *W, int k, int i, const uint32_t *s, double *phi, double *tanh_)
{
    prob_t p = mkprob(N, M, K, 0.0, NULL);
    cplx *scr = (cplx *)malloc(sizeof(cplx) * M);
    flip_row(&p, (const cplx *)W, k, i, s, (cplx *)phi, (cplx *)tanh_, scr);
    free(scr);
}

double orc_ezz(int N, const double *J, const uint32_t *s)
{
    prob_t p = mkprob(N, 1, 1, 0.0, J);
    return ezz(&p, s);
}

void orc_hloc_model(int model, int N, int M, int K, double h, const double *J, const double *W, const uint32_t *s,
                    const double *phi, const double *phi_inv, const double *tanh_, double *H)
{
    prob_t p = mkprob(N, M, K, h, J);
    p.model = model;
    cplx *scr = (cplx *)malloc(sizeof(cplx) * 2 * M);
    hloc(&p, (const cplx *)W, s, (const cplx *)phi, (const cplx *)phi_inv, (const cplx *)tanh_, (cplx *)H, scr);
    free(scr);
}

void orc_hloc(int N, int M, int K, double h, const double *J, const double *W, const uint32_t *s,
              const double *phi, const double *phi_inv, const double *tanh_, double *H)
{
    orc_hloc_model(0, N, M, K, h, J, W, s, phi, phi_inv, tanh_, H);
}

void orc_deriv(int N, int M, int K, const uint32_t *s, const double *phi, const double *phi_inv,
               const double *tanh_, double *O)
{
    prob_t p = mkprob(N, M, K, 0.0, NULL);
    deriv(&p, s, (const cplx *)phi, (const cplx *)phi_inv, (const cplx *)tanh_, (cplx *)O);
}

/* ------------------------------------------------------------------------- */
/* Metropolis sweep (rbm.cc:530-552)                                         */
/* ------------------------------------------------------------------------- */
typedef struct {
    cplx *phi, *phi_inv, *tanh_;   /* current state */
    cplx *phip, *tanhp;            /* proposal copies */
    cplx *scratch;
    cplx det;
    long accepts;
} chain_t;

static void chain_alloc(const prob_t *p, chain_t *c)
{
    int K = p->K, M = p->M;
    c->phi = (cplx *)malloc(sizeof(cplx) * K * K);
    c->phi_inv = (cplx *)malloc(sizeof(cplx) * K * K);
    c->phip = (cplx *)malloc(sizeof(cplx) * K * K);
    c->tanh_ = (cplx *)malloc(sizeof(cplx) * M * K * K);
    c->tanhp = (cplx *)malloc(sizeof(cplx) * M * K * K);
    c->scratch = (cplx *)malloc(sizeof(cplx) * 2 * M);
    c->accepts = 0;
}

static void chain_free(chain_t *c)
{
    free(c->phi); free(c->phi_inv); free(c->phip); free(c->tanh_); free(c->tanhp); free(c->scratch);
}

static void mc_sweep(const prob_t *p, const cplx *W, uint32_t *s, u128 *rng, chain_t *c)
{
    int K = p->K, M = p->M;
    cplx det_phi = detK(c->phi, K);
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < p->N; i++) {
            memcpy(c->phip, c->phi, sizeof(cplx) * K * K);
            /* only the K columns (k, kk) of tanh change: copy those */
            memcpy(c->tanhp + (long)k * K * M, c->tanh_ + (long)k * K * M, sizeof(cplx) * K * M);
            flip_row(p, W, k, i, s, c->phip, c->tanhp, c->scratch);
            cplx det_p = detK(c->phip, K);
            double prob = (det_phi.re == 0.0 && det_phi.im == 0.0) ? 1.0 : cnorm(det_p) / cnorm(det_phi);
            double u = canonical(pcg_next(rng));
            if (u <= (prob < 1.0 ? prob : 1.0)) {
                sflip(s + k * p->NW, i);
                det_phi = det_p;
                memcpy(c->phi, c->phip, sizeof(cplx) * K * K);
                invK(c->phi, K, c->phi_inv);
                memcpy(c->tanh_ + (long)k * K * M, c->tanhp + (long)k * K * M, sizeof(cplx) * K * M);
                c->accepts++;
            }
        }
    }
    c->det = det_phi;
}

/* chain start: spins, state_fcn, inverse (rbm.cc:604-608) */
static void chain_start(const prob_t *p, const cplx *W, uint32_t *s, u128 *rng, chain_t *c)
{
    init_spins(p, rng, s);
    state_fcn(p, W, s, c->phi, c->tanh_);
    invK(c->phi, p->K, c->phi_inv);
}

/*
 * Runs chain c of one SR step (rbm.cc:603-637): rng = pcg64(seed + c), spins,
 * B burn-in sweeps, then m sweeps each followed by H_loc and O.  Writes per
 * sample: spins (K*NW words), H (K*K complex), optionally O (P complex) and the
 * state (phi, phi_inv: K*K each; tanh: M*K*K).  Any output may be NULL.
 */
void orc_run_chain_model(int model, int N, int M, int K, double h, const double *J, const double *Wd,
                         uint64_t seed, int c, int m, int B,
                         uint32_t *spins_out, double *H_out, double *O_out,
                         double *phi_out, double *phi_inv_out, double *tanh_out, long *accepts)
{
    prob_t p = mkprob(N, M, K, h, J);
    p.model = model;
    const cplx *W = (const cplx *)Wd;
    u128 rng = pcg_seed(seed + (uint64_t)c);
    uint32_t s[ORC_KMAX * (ORC_NMAX / 32)];
    chain_t ch;
    chain_alloc(&p, &ch);
    chain_start(&p, W, s, &rng, &ch);
    for (int t = 0; t < B; t++) mc_sweep(&p, W, s, &rng, &ch);
    for (int t = 0; t < m; t++) {
        mc_sweep(&p, W, s, &rng, &ch);
        if (spins_out) memcpy(spins_out + (long)t * K * p.NW, s, sizeof(uint32_t) * K * p.NW);
        if (H_out) hloc(&p, W, s, ch.phi, ch.phi_inv, ch.tanh_, (cplx *)H_out + (long)t * K * K, ch.scratch);
        if (O_out) deriv(&p, s, ch.phi, ch.phi_inv, ch.tanh_, (cplx *)O_out + (long)t * p.P);
        if (phi_out) memcpy((cplx *)phi_out + (long)t * K * K, ch.phi, sizeof(cplx) * K * K);
        if (phi_inv_out) memcpy((cplx *)phi_inv_out + (long)t * K * K, ch.phi_inv, sizeof(cplx) * K * K);
        if (tanh_out) memcpy((cplx *)tanh_out + (long)t * M * K * K, ch.tanh_, sizeof(cplx) * M * K * K);
    }
    if (accepts) *accepts = ch.accepts;
    chain_free(&ch);
}

/*
 * Chains [0, C) of one SR step in parallel (OpenMP over chains), each as
 * orc_run_chain_model: spins (C*m*K*NW words), H (C*m*K*K complex), optionally O
 * (C*m*P complex, NULL: not formed) and per-chain accept counts (C, may be NULL).
 */
void orc_sample_chains_model(int model, int N, int M, int K, double h, const double *J, const double *Wd,
                             uint64_t seed, int C, int m, int B, uint32_t *spins_out, double *H_out,
                             double *O_out, long *accepts)
{
    const int NW = (N + 31) / 32;
    const long P = (long)K * (N + M + (long)N * M);
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < C; c++) {
        long acc = 0;
        orc_run_chain_model(model, N, M, K, h, J, Wd, seed, c, m, B,
                            spins_out ? spins_out + (long)c * m * K * NW : NULL,
                            H_out ? H_out + 2L * c * m * K * K : NULL,
                            O_out ? O_out + 2L * c * m * P : NULL, NULL, NULL, NULL, &acc);
        if (accepts) accepts[c] = acc;
    }
}

void orc_run_chain(int N, int M, int K, double h, const double *J, const double *Wd,
                   uint64_t seed, int c, int m, int B,
                   uint32_t *spins_out, double *H_out, double *O_out,
                   double *phi_out, double *phi_inv_out, double *tanh_out, long *accepts)
{
    orc_run_chain_model(0, N, M, K, h, J, Wd, seed, c, m, B, spins_out, H_out, O_out, phi_out, phi_inv_out,
                        tanh_out, accepts);
}

/*
 * correlation_fcn (rbm.cc:730-771) for the four operators of rbm.cc:774-829, all
 * accumulated over one set of chains (the reference runs the same chains four
 * times with the same seed).  sigma^x terms recompute Phi from scratch with
 * state_fcn on the flipped configuration, as corrfcn_x / corrfcn_xx do.  Outputs
 * the sample means of Phi^+ fcn(idx): x[N], z[N], xx[N*N], zz[N*N] (K*K complex each).
 */
void orc_correlations(int N, int M, int K, double h, const double *J, const double *Wd, uint64_t seed,
                      int C, int m, int B, double *x, double *z, double *xx, double *zz)
{
    prob_t p = mkprob(N, M, K, h, J);
    const cplx *W = (const cplx *)Wd;
    long KK = (long)K * K;
    long ncor = 2L * N + 2L * N * N;
    cplx *cor = (cplx *)calloc((size_t)ncor * KK, sizeof(cplx));
    for (int c = 0; c < C; c++) {
        u128 rng = pcg_seed(seed + (uint64_t)c);
        uint32_t s[ORC_KMAX * (ORC_NMAX / 32)], s2[ORC_KMAX * (ORC_NMAX / 32)];
        chain_t ch;
        chain_alloc(&p, &ch);
        chain_start(&p, W, s, &rng, &ch);
        for (int t = 0; t < B; t++) mc_sweep(&p, W, s, &rng, &ch);
        for (int t = 0; t < m; t++) {
            mc_sweep(&p, W, s, &rng, &ch);
            cplx X[ORC_KMAX * ORC_KMAX], mat[ORC_KMAX * ORC_KMAX];
            for (long idx = 0; idx < ncor; idx++) {
                int i, j;
                if (idx < N) {                                /* sigma^x_i */
                    i = (int)idx;
                    memcpy(s2, s, sizeof(uint32_t) * K * p.NW);
                    for (int k = 0; k < K; k++) sflip(s2 + k * p.NW, i);
                    state_fcn(&p, W, s2, X, NULL);
                } else if (idx < 2L * N) {                    /* sigma^z_i */
                    i = (int)(idx - N);
                    for (int k = 0; k < K; k++)
                        for (int kk = 0; kk < K; kk++) {
                            cplx v = ch.phi[k * K + kk];
                            X[k * K + kk] = sbit(s + k * p.NW, i) ? cneg(v) : v;
                        }
                } else if (idx < 2L * N + (long)N * N) {      /* sigma^x_i sigma^x_j */
                    long r = idx - 2L * N;
                    i = (int)(r / N);
                    j = (int)(r % N);
                    memcpy(s2, s, sizeof(uint32_t) * K * p.NW);
                    for (int k = 0; k < K; k++) {
                        sflip(s2 + k * p.NW, i);
                        sflip(s2 + k * p.NW, j);
                    }
                    state_fcn(&p, W, s2, X, NULL);
                } else {                                      /* sigma^z_i sigma^z_j */
                    long r = idx - 2L * N - (long)N * N;
                    i = (int)(r / N);
                    j = (int)(r % N);
                    for (int k = 0; k < K; k++)
                        for (int kk = 0; kk < K; kk++) {
                            cplx v = ch.phi[k * K + kk];
                            int f = sbit(s + k * p.NW, i) ^ sbit(s + k * p.NW, j);
                            X[k * K + kk] = f ? cneg(v) : v;
                        }
                }
                matmulK(ch.phi_inv, X, K, mat);
                for (long q = 0; q < KK; q++) cor[idx * KK + q] = cadd(cor[idx * KK + q], mat[q]);
            }
        }
        chain_free(&ch);
    }
    double Sd = (double)C * (double)m;
    double *outs[4] = {x, z, xx, zz};
    long off[4] = {0, N, 2L * N, 2L * N + (long)N * N};
    long cnt[4] = {N, N, (long)N * N, (long)N * N};
    for (int t = 0; t < 4; t++)
        for (long q = 0; q < cnt[t] * KK; q++) {
            cplx v = cdivr(cor[off[t] * KK + q], Sd);
            outs[t][2 * q] = v.re;
            outs[t][2 * q + 1] = v.im;
        }
    free(cor);
}

/* ------------------------------------------------------------------------- */
/* MINRES-QLP (minresqlp.h:85-390), complex vectors, real scalars             */
/* ------------------------------------------------------------------------- */
typedef void (*orc_op_t)(void *ctx, const cplx *v, cplx *out);

static double sgn(double a) { return copysign(1.0, a); }

/* sym_ortho (minresqlp.h:85-110) */
static void sym_ortho(double a, double b, double *c, double *s, double *r)
{
    if (b == 0.0) {
        *c = (a == 0.0) ? 1.0 : sgn(a);
        *s = 0.0;
        *r = fabs(a);
    } else if (a == 0.0) {
        *c = 0.0;
        *s = sgn(b);
        *r = fabs(b);
    } else if (fabs(b) > fabs(a)) {
        double t = a / b;
        *s = sgn(b) / sqrt(1.0 + t * t);
        *c = *s * t;
        *r = b / *s;
    } else {
        double t = b / a;
        *c = sgn(a) / sqrt(1.0 + t * t);
        *s = *c * t;
        *r = a / *c;
    }
}

static double vnorm(const cplx *x, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += x[i].re * x[i].re + x[i].im * x[i].im;
    return sqrt(s);
}

/* Re(x^H y) */
static double vdot_re(const cplx *x, const cplx *y, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; i++) s += x[i].re * y[i].re + x[i].im * y[i].im;
    return s;
}

enum { MQ_INTERNAL = -2, MQ_EIGVEC = -1, MQ_ZERO = 0, MQ_CONV = 1, MQ_CONV_LS = 2,
       MQ_EPS = 3, MQ_EPS_LS = 4, MQ_EIGVEC_CONV = 5, MQ_XNORM = 6, MQ_ACOND = 7,
       MQ_MAXIT = 8, MQ_SINGULAR = 9 };

/* Solves A x = b (A Hermitian, given as an operator) by MINRES-QLP with
 * shift = 0, maxxnorm = 1e7, Acondlim = 1e15, TranCond = 1e7 (the defaults the
 * reference's call sites use, rbm.cc:661, 686).  Returns the flag. */
static int minresqlp(orc_op_t A, void *ctx, const cplx *b, cplx *x, long n, double rtol,
                     uint64_t maxit, uint64_t *iters_out)
{
    const double maxxnorm = 1e7, Acondlim = 1e15, TranCond = 1e7;
    size_t bytes = sizeof(cplx) * (size_t)n;
    cplx *r1 = (cplx *)malloc(bytes), *r2 = (cplx *)malloc(bytes), *r3 = (cplx *)malloc(bytes);
    cplx *v = (cplx *)malloc(bytes), *w = (cplx *)calloc(n, sizeof(cplx));
    cplx *wl = (cplx *)calloc(n, sizeof(cplx)), *wl2 = (cplx *)calloc(n, sizeof(cplx));
    cplx *xl2 = (cplx *)calloc(n, sizeof(cplx)), *tmp = (cplx *)malloc(bytes);
    memcpy(r1, b, bytes); memcpy(r2, b, bytes); memcpy(r3, b, bytes);
    double beta1 = vnorm(r2, n);

    int flag0 = MQ_INTERNAL, flag = flag0;
    uint64_t iters = 0, QLPiter = 0;
    double beta = 0, tau = 0, taul = 0, phi = beta1, betan = beta1, gmin = 0, gminl = 0;
    double cs = -1, sn = 0, cr1 = -1, sr1 = 0, cr2 = -1, sr2 = 0, dltan = 0, eplnn = 0;
    double gama = 0, gamal = 0, gamal2 = 0, eta = 0, etal = 0, etal2 = 0;
    double vepln = 0, veplnl = 0, veplnl2 = 0, ul3 = 0, ul2 = 0, ul = 0, u = 0;
    double gama_QLP = 0, gamal_QLP = 0, vepln_QLP = 0, u_QLP = 0, ul_QLP = 0;
    double rnorm = betan, xnorm = 0, xl2norm = 0, Axnorm = 0, Anorm = 0, Acond = 1;
    double relres = rnorm / (beta1 + 1e-50);
    for (long i = 0; i < n; i++) x[i] = cmk(0.0, 0.0);
    if (beta1 == 0.0) flag = MQ_ZERO;

    while (flag == flag0 && iters < maxit) {
        /* Lanczos step */
        iters++;
        double betal = beta;
        beta = betan;
        for (long i = 0; i < n; i++) v[i] = cdivr(r3[i], beta);
        A(ctx, v, r3);
        if (iters > 1) {
            double f = beta / betal;
            for (long i = 0; i < n; i++) r3[i] = csub(r3[i], cscale(f, r1[i]));
        }
        double alfa = vdot_re(r3, v, n);
        {
            double f = alfa / beta;
            for (long i = 0; i < n; i++) r3[i] = csub(r3[i], cscale(f, r2[i]));
        }
        memcpy(r1, r2, bytes);
        memcpy(r2, r3, bytes);
        betan = vnorm(r3, n);
        if (iters == 1 && betan == 0.0) {
            if (alfa == 0.0) {
                flag = MQ_ZERO;
            } else {
                flag = MQ_EIGVEC;
                for (long i = 0; i < n; i++) x[i] = cdivr(b[i], alfa);
            }
            break;
        }
        double pnorm = sqrt(betal * betal + alfa * alfa + betan * betan);

        /* previous left rotation Q_{k-1} */
        double dbar = dltan;
        double dlta = cs * dbar + sn * alfa;
        double epln = eplnn;
        double gbar = sn * dbar - cs * alfa;
        eplnn = sn * betan;
        dltan = -cs * betan;
        double dlta_QLP = dlta;
        /* current left plane rotation Q_k */
        double gamal3 = gamal2;
        gamal2 = gamal;
        gamal = gama;
        sym_ortho(gbar, betan, &cs, &sn, &gama);
        double gama_tmp = gama;
        double taul2 = taul;
        taul = tau;
        tau = cs * phi;
        Axnorm = sqrt(Axnorm * Axnorm + tau * tau);
        phi = sn * phi;
        /* previous right plane rotation P_{k-2,k} */
        if (iters > 2) {
            veplnl2 = veplnl;
            etal2 = etal;
            etal = eta;
            double dlta_tmp = sr2 * vepln - cr2 * dlta;
            veplnl = cr2 * vepln + sr2 * dlta;
            dlta = dlta_tmp;
            eta = sr2 * gama;
            gama = -cr2 * gama;
        }
        /* current right plane rotation P_{k-1,k} */
        if (iters > 1) {
            sym_ortho(gamal, dlta, &cr1, &sr1, &gamal);
            vepln = sr1 * gama;
            gama = -cr1 * gama;
        }

        /* xnorm update */
        double ul4 = ul3;
        ul3 = ul2;
        if (iters > 2) ul2 = (taul2 - etal2 * ul4 - veplnl2 * ul3) / gamal2;
        if (iters > 1) ul = (taul - etal * ul3 - veplnl * ul2) / gamal;
        double xnorm_tmp = sqrt(xl2norm * xl2norm + ul2 * ul2 + ul * ul);
        if (fabs(gama) > DBL_MIN && xnorm_tmp < maxxnorm) {
            u = (tau - eta * ul2 - vepln * ul) / gama;
            if (sqrt(xnorm_tmp * xnorm_tmp + u * u) > maxxnorm) {
                u = 0;
                flag = MQ_XNORM;
            }
        } else {
            u = 0;
            flag = MQ_SINGULAR;
        }
        xl2norm = sqrt(xl2norm * xl2norm + ul2 * ul2);
        xnorm = sqrt(xl2norm * xl2norm + ul * ul + u * u);

        /* update w and x */
        if (Acond < TranCond && flag != flag0 && QLPiter == 0) {
            /* MINRES step */
            memcpy(wl2, wl, bytes);
            memcpy(wl, w, bytes);
            for (long i = 0; i < n; i++)
                w[i] = cdivr(csub(csub(v[i], cscale(epln, wl2[i])), cscale(dlta_QLP, wl[i])), gama_tmp);
            if (xnorm < maxxnorm) {
                for (long i = 0; i < n; i++) x[i] = cadd(x[i], cscale(tau, w[i]));
            } else {
                flag = MQ_XNORM;
            }
        } else {
            /* MINRES-QLP step */
            QLPiter++;
            if (QLPiter == 1 && iters > 1) {
                /* rebuild w_{k-3}, w_{k-2}, w_{k-1} */
                if (iters > 3)
                    for (long i = 0; i < n; i++)
                        wl2[i] = cadd(cadd(cscale(gamal3, wl2[i]), cscale(veplnl2, wl[i])), cscale(etal, w[i]));
                if (iters > 2)
                    for (long i = 0; i < n; i++)
                        wl[i] = cadd(cscale(gamal_QLP, wl[i]), cscale(vepln_QLP, w[i]));
                for (long i = 0; i < n; i++) w[i] = cscale(gama_QLP, w[i]);
                for (long i = 0; i < n; i++)
                    xl2[i] = csub(csub(x[i], cscale(ul_QLP, wl[i])), cscale(u_QLP, w[i]));
            }
            if (iters == 1) {
                memcpy(wl2, wl, bytes);
                for (long i = 0; i < n; i++) wl[i] = cscale(sr1, v[i]);
                for (long i = 0; i < n; i++) w[i] = cscale(cr1, cneg(v[i]));
            } else if (iters == 2) {
                memcpy(wl2, wl, bytes);
                for (long i = 0; i < n; i++) {
                    cplx wi = w[i], vi = v[i];
                    wl[i] = cadd(cscale(cr1, wi), cscale(sr1, vi));
                    w[i] = csub(cscale(sr1, wi), cscale(cr1, vi));
                }
            } else {
                for (long i = 0; i < n; i++) {
                    cplx o_wl = wl[i], o_w = w[i], vi = v[i];
                    /* wl2 <- wl; wl <- w; w <- wl2*sr2 - v*cr2; wl2 <- wl2*cr2 + v*sr2;
                     * v <- wl*cr1 + w*sr1; w <- wl*sr1 - w*cr1; wl <- v */
                    cplx nw = csub(cscale(sr2, o_wl), cscale(cr2, vi));
                    cplx nwl2 = cadd(cscale(cr2, o_wl), cscale(sr2, vi));
                    cplx nv = cadd(cscale(cr1, o_w), cscale(sr1, nw));
                    cplx nw2 = csub(cscale(sr1, o_w), cscale(cr1, nw));
                    wl2[i] = nwl2;
                    v[i] = nv;
                    w[i] = nw2;
                    wl[i] = nv;
                }
            }
            for (long i = 0; i < n; i++) xl2[i] = cadd(xl2[i], cscale(ul2, wl2[i]));
            for (long i = 0; i < n; i++)
                x[i] = cadd(cadd(xl2[i], cscale(ul, wl[i])), cscale(u, w[i]));
        }

        /* next right plane rotation P_{k-1,k+1} */
        double gamal_tmp = gamal;
        sym_ortho(gamal, eplnn, &cr2, &sr2, &gamal);
        gamal_QLP = gamal_tmp;
        vepln_QLP = vepln;
        gama_QLP = gama;
        ul_QLP = ul;
        u_QLP = u;
        /* norm estimates */
        double abs_gama = fabs(gama);
        Anorm = fmax(fmax(Anorm, pnorm), fmax(gamal, abs_gama));
        if (iters == 1) {
            gmin = gama;
            gminl = gmin;
        } else {
            double gminl2 = gminl;
            gminl = gmin;
            gmin = fmin(fmin(gminl2, gamal), abs_gama);
        }
        double Acondl = Acond;
        Acond = Anorm / gmin;
        double rnorml = rnorm;
        double relresl = relres;
        if (flag != MQ_SINGULAR) rnorm = phi;
        relres = rnorm / (Anorm * xnorm + beta1);
        double rootl = sqrt(gbar * gbar + dltan * dltan);
        double relAresl = rootl / Anorm;
        double epsx = Anorm * xnorm * DBL_EPSILON;
        if (flag == flag0 || flag == MQ_SINGULAR) {
            double t1 = 1 + relres, t2 = 1 + relAresl;
            if (iters >= maxit) flag = MQ_MAXIT;
            if (Acond >= Acondlim) flag = MQ_ACOND;
            if (xnorm >= maxxnorm) flag = MQ_XNORM;
            if (epsx >= beta1) flag = MQ_EIGVEC_CONV;
            if (t2 <= 1) flag = MQ_EPS_LS;
            if (t1 <= 1) flag = MQ_EPS;
            if (relAresl <= rtol) flag = MQ_CONV_LS;
            if (relres <= rtol) flag = MQ_CONV;
        }
        if (flag == MQ_CONV_LS || flag == MQ_EPS_LS || flag == MQ_XNORM || flag == MQ_ACOND) {
            iters--;
            Acond = Acondl;
            rnorm = rnorml;
            relres = relresl;
        }
    }
    if (iters_out) *iters_out = iters;
    free(r1); free(r2); free(r3); free(v); free(w); free(wl); free(wl2); free(xl2); free(tmp);
    return flag;
}

/* dense Hermitian operator for tests */
typedef struct { long n; const cplx *A; } dense_ctx;
static void dense_apply(void *c, const cplx *v, cplx *out)
{
    dense_ctx *d = (dense_ctx *)c;
    for (long r = 0; r < d->n; r++) {
        cplx acc = cmk(0.0, 0.0);
        for (long q = 0; q < d->n; q++) acc = cadd(acc, cmul(d->A[r * d->n + q], v[q]));
        out[r] = acc;
    }
}

int orc_minresqlp_dense(long n, const double *A, const double *b, double *x, double rtol,
                        uint64_t maxit, uint64_t *iters)
{
    dense_ctx d = { n, (const cplx *)A };
    return minresqlp(dense_apply, &d, (const cplx *)b, (cplx *)x, n, rtol, maxit, iters);
}

/* ------------------------------------------------------------------------- */
/* SR step (monte_carlo, rbm.cc:564-720), matrix-free                        */
/* ------------------------------------------------------------------------- */
#define ORC_NBLK 64   /* fixed chain blocks: reductions are independent of thread count */

typedef struct {
    prob_t p;
    long S;
    int C, m;
    uint32_t *spins;             /* S x K*NW  */
    cplx *phis, *phi_invs, *tanhs;
    const cplx *o, *skk;
    cplx **blk_res;              /* ORC_NBLK partial results */
    cplx *O_store;               /* optional S x P stored O (NULL: recompute) */
} sv_ctx;

/* S*v lambda (rbm.cc:661-684):
 * res = (1/S) sum_s (O_s^T v) conj(O_s) + skk .* v - (o^T v) conj(o) */
static void sv_apply(void *vc, const cplx *v, cplx *out)
{
    sv_ctx *c = (sv_ctx *)vc;
    const prob_t *p = &c->p;
    long P = p->P;
    int K = p->K, M = p->M;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < ORC_NBLK; b++) {
        cplx *res = c->blk_res[b];
        cplx *O = c->O_store ? NULL : (cplx *)malloc(sizeof(cplx) * P);
        for (long q = 0; q < P; q++) res[q] = cmk(0.0, 0.0);
        long c0 = (long)c->C * b / ORC_NBLK, c1 = (long)c->C * (b + 1) / ORC_NBLK;
        for (long s = c0 * c->m; s < c1 * c->m; s++) {
            const cplx *Os;
            if (c->O_store) {
                Os = c->O_store + s * P;
            } else {
                deriv(p, c->spins + s * K * p->NW, c->phis + s * K * K, c->phi_invs + s * K * K,
                      c->tanhs + s * (long)M * K * K, O);
                Os = O;
            }
            cplx ov = cmk(0.0, 0.0);
            for (long q = 0; q < P; q++) ov = cadd(ov, cmul(Os[q], v[q]));
            for (long q = 0; q < P; q++) res[q] = cadd(res[q], cmul(ov, cconj(Os[q])));
        }
        free(O);
    }
    for (long q = 0; q < P; q++) {
        cplx acc = c->blk_res[0][q];
        for (int b = 1; b < ORC_NBLK; b++) acc = cadd(acc, c->blk_res[b][q]);
        out[q] = acc;
    }
    cplx ov = cmk(0.0, 0.0);
    for (long q = 0; q < P; q++) ov = cadd(ov, cmul(c->o[q], v[q]));
    for (long q = 0; q < P; q++) {
        out[q] = cdivr(out[q], (double)c->S);
        out[q] = cadd(out[q], csub(cmul(c->skk[q], v[q]), cmul(ov, cconj(c->o[q]))));
    }
}

/*
 * One SR step.  W (P complex) is updated in place; emat (K*K complex) receives
 * e_mat / S.  If store_O != 0 the O rows are kept instead of recomputed inside
 * S*v (same arithmetic).  Optional outputs: x (P complex: the SR solution),
 * o/skk/f (P complex each, after normalisation), eloc (1 complex).
 * Returns the MINRES-QLP flag.
 */
int orc_sr_step_model(int model, int N, int M, int K, double h, const double *J, double *Wd,
                      uint64_t seed, int C, int m, int B, double lr, double reglam,
                      double rtol, uint64_t maxit, int store_O,
                      double *emat_out, uint64_t *iters_out, long *accepts_out,
                      double *x_out, double *o_out, double *skk_out, double *f_out, double *eloc_out)
{
    prob_t p = mkprob(N, M, K, h, J);
    p.model = model;
    cplx *W = (cplx *)Wd;
    long P = p.P, S = (long)C * m;
    int NW = p.NW;
    sv_ctx sv;
    sv.p = p; sv.S = S; sv.C = C; sv.m = m;
    sv.spins = (uint32_t *)malloc(sizeof(uint32_t) * S * K * NW);
    sv.phis = (cplx *)malloc(sizeof(cplx) * S * K * K);
    sv.phi_invs = (cplx *)malloc(sizeof(cplx) * S * K * K);
    sv.tanhs = (cplx *)malloc(sizeof(cplx) * S * (long)M * K * K);
    sv.O_store = store_O ? (cplx *)malloc(sizeof(cplx) * S * P) : NULL;
    /* per-block accumulators */
    cplx *bo = (cplx *)calloc((size_t)ORC_NBLK * P, sizeof(cplx));
    cplx *bf = (cplx *)calloc((size_t)ORC_NBLK * P, sizeof(cplx));
    cplx *bs = (cplx *)calloc((size_t)ORC_NBLK * P, sizeof(cplx));
    cplx be[ORC_NBLK][ORC_KMAX * ORC_KMAX];
    cplx bel[ORC_NBLK];
    long bacc[ORC_NBLK];

#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < ORC_NBLK; b++) {
        cplx *o = bo + (long)b * P, *f = bf + (long)b * P, *sk = bs + (long)b * P;
        cplx *O = (cplx *)malloc(sizeof(cplx) * P);
        cplx H[ORC_KMAX * ORC_KMAX];
        for (int q = 0; q < K * K; q++) be[b][q] = cmk(0.0, 0.0);
        bel[b] = cmk(0.0, 0.0);
        bacc[b] = 0;
        chain_t ch;
        chain_alloc(&p, &ch);
        long c0 = (long)C * b / ORC_NBLK, c1 = (long)C * (b + 1) / ORC_NBLK;
        for (long c = c0; c < c1; c++) {
            u128 rng = pcg_seed(seed + (uint64_t)c);
            uint32_t s[ORC_KMAX * (ORC_NMAX / 32)];
            ch.accepts = 0;
            chain_start(&p, W, s, &rng, &ch);
            for (int t = 0; t < B; t++) mc_sweep(&p, W, s, &rng, &ch);
            for (int t = 0; t < m; t++) {
                mc_sweep(&p, W, s, &rng, &ch);
                hloc(&p, W, s, ch.phi, ch.phi_inv, ch.tanh_, H, ch.scratch);
                cplx energy = cmk(0.0, 0.0);
                for (int q = 0; q < K * K; q++) be[b][q] = cadd(be[b][q], H[q]);
                for (int k = 0; k < K; k++) energy = cadd(energy, H[k * K + k]);
                bel[b] = cadd(bel[b], energy);
                deriv(&p, s, ch.phi, ch.phi_inv, ch.tanh_, O);
                for (long q = 0; q < P; q++) {
                    o[q] = cadd(o[q], O[q]);
                    sk[q] = cadd(sk[q], cmul(O[q], cconj(O[q])));
                    f[q] = cadd(f[q], cmul(energy, cconj(O[q])));
                }
                long idx = c * m + t;
                memcpy(sv.spins + idx * K * NW, s, sizeof(uint32_t) * K * NW);
                memcpy(sv.phis + idx * K * K, ch.phi, sizeof(cplx) * K * K);
                memcpy(sv.phi_invs + idx * K * K, ch.phi_inv, sizeof(cplx) * K * K);
                memcpy(sv.tanhs + idx * (long)M * K * K, ch.tanh_, sizeof(cplx) * M * K * K);
                if (sv.O_store) memcpy(sv.O_store + idx * P, O, sizeof(cplx) * P);
            }
            bacc[b] += ch.accepts;
        }
        chain_free(&ch);
        free(O);
    }
    /* combine blocks in order */
    cplx *o = (cplx *)malloc(sizeof(cplx) * P), *f = (cplx *)malloc(sizeof(cplx) * P);
    cplx *skk = (cplx *)malloc(sizeof(cplx) * P);
    cplx emat[ORC_KMAX * ORC_KMAX], eloc = bel[0];
    long acc = 0;
    for (int q = 0; q < K * K; q++) emat[q] = be[0][q];
    for (int b = 0; b < ORC_NBLK; b++) acc += bacc[b];
    for (int b = 1; b < ORC_NBLK; b++) {
        eloc = cadd(eloc, bel[b]);
        for (int q = 0; q < K * K; q++) emat[q] = cadd(emat[q], be[b][q]);
    }
    for (long q = 0; q < P; q++) {
        cplx so = bo[q], sf = bf[q], ss = bs[q];
        for (int b = 1; b < ORC_NBLK; b++) {
            so = cadd(so, bo[(long)b * P + q]);
            sf = cadd(sf, bf[(long)b * P + q]);
            ss = cadd(ss, bs[(long)b * P + q]);
        }
        o[q] = so; f[q] = sf; skk[q] = ss;
    }
    free(bo); free(bf); free(bs);

    /* normalisation (rbm.cc:641-650) */
    double Sd = (double)S;
    eloc = cdivr(eloc, Sd);
    for (long q = 0; q < P; q++) {
        o[q] = cdivr(o[q], Sd);
        skk[q] = cdivr(skk[q], Sd);
        skk[q].re -= cnorm(o[q]);
        skk[q] = cscale(reglam, skk[q]);
        f[q] = cdivr(f[q], Sd);
        f[q] = csub(f[q], cmul(eloc, cconj(o[q])));
    }
    if (o_out) memcpy(o_out, o, sizeof(cplx) * P);
    if (skk_out) memcpy(skk_out, skk, sizeof(cplx) * P);
    if (f_out) memcpy(f_out, f, sizeof(cplx) * P);
    if (eloc_out) { eloc_out[0] = eloc.re; eloc_out[1] = eloc.im; }

    sv.o = o; sv.skk = skk;
    sv.blk_res = (cplx **)malloc(sizeof(cplx *) * ORC_NBLK);
    for (int b = 0; b < ORC_NBLK; b++) sv.blk_res[b] = (cplx *)malloc(sizeof(cplx) * P);
    cplx *x = (cplx *)malloc(sizeof(cplx) * P);
    uint64_t iters = 0;
    int flag = minresqlp(sv_apply, &sv, f, x, P, rtol, maxit, &iters);
    if (x_out) memcpy(x_out, x, sizeof(cplx) * P);

    /* weight update w -= lr * x (rbm.cc:699-717) */
    for (long q = 0; q < P; q++) W[q] = csub(W[q], cscale(lr, x[q]));
    for (int q = 0; q < K * K; q++) {
        cplx e = cdivr(emat[q], Sd);
        emat_out[2 * q] = e.re; emat_out[2 * q + 1] = e.im;
    }
    if (iters_out) *iters_out = iters;
    if (accepts_out) *accepts_out = acc;

    for (int b = 0; b < ORC_NBLK; b++) free(sv.blk_res[b]);
    free(sv.blk_res);
    free(x); free(o); free(f); free(skk);
    free(sv.spins); free(sv.phis); free(sv.phi_invs); free(sv.tanhs); free(sv.O_store);
    return flag;
}

int orc_sr_step(int N, int M, int K, double h, const double *J, double *Wd,
                uint64_t seed, int C, int m, int B, double lr, double reglam,
                double rtol, uint64_t maxit, int store_O,
                double *emat_out, uint64_t *iters_out, long *accepts_out,
                double *x_out, double *o_out, double *skk_out, double *f_out, double *eloc_out)
{
    return orc_sr_step_model(0, N, M, K, h, J, Wd, seed, C, m, B, lr, reglam, rtol, maxit, store_O, emat_out,
                             iters_out, accepts_out, x_out, o_out, skk_out, f_out, eloc_out);
}

/* S*v on a given set of stored O rows (tests: matrix-free vs dense S) */
int orc_sv_from_O(long S, long P, const double *O, const double *o, const double *skk,
                  const double *v, double *out)
{
    sv_ctx sv;
    memset(&sv, 0, sizeof(sv));
    sv.p.P = P; sv.p.K = 1; sv.p.M = 1; sv.p.NW = 1;
    sv.S = S; sv.C = (int)S; sv.m = 1;
    sv.O_store = (cplx *)O;
    sv.o = (const cplx *)o; sv.skk = (const cplx *)skk;
    sv.blk_res = (cplx **)malloc(sizeof(cplx *) * ORC_NBLK);
    for (int b = 0; b < ORC_NBLK; b++) sv.blk_res[b] = (cplx *)malloc(sizeof(cplx) * P);
    sv_apply(&sv, (const cplx *)v, (cplx *)out);
    for (int b = 0; b < ORC_NBLK; b++) free(sv.blk_res[b]);
    free(sv.blk_res);
    return 0;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
