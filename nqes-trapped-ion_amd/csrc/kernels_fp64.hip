// kernels_fp64.hip -- FP64 device path of one SR step.
//
// Compiled with -ffp-contract=off: the arithmetic follows the canonical order of
// cplx64.hpp, so spins, H_loc and O agree bit for bit with any implementation of
// the same order (the CPU oracle is one).
//
// Kernels
//   k_tables64      cosh(2w), sinh(2w) per (RBM kk, spin i, hidden j) and
//                   exp(+-2a) per (kk, i), once per SR step (weights fixed within
//                   a step; rbm.cc:181-186 evaluates them per proposal).
//   k_sample64      one wavefront per Markov chain: chain start (rbm.cc:604-608),
//                   B burn-in + m sampled Metropolis sweeps (rbm.cc:530-552), and
//                   after each sampled sweep the local energy H_loc
//                   (rbm.cc:1042-1064) and the log-derivative row O (rbm.cc:199-239)
//                   written coalesced to HBM.
//   k_moments64     column sums of O: sum O, sum O conj(O), sum E conj(O)
//                   (rbm.cc:620-626), two-stage and deterministic.
//   k_colsum64      column sums of a small S x n complex matrix (H_loc, E_loc).
//   k_sv_rowdot64   w_s = O_s^T v                (first pass of S*v, rbm.cc:668)
//   k_sv_colacc64   r_p = sum_s w_s conj(O_sp)   (second pass,        rbm.cc:668)
//
// Lane layout of k_sample64: lane g owns hidden units j = g + 64 q (q < JPL) for all
// K RBMs.  A product over j (rbm.cc:159-168, 189-194) is a per-lane product over q in
// ascending j followed by an xor-butterfly over the 64 lanes; since complex multiply
// is commutative bit for bit, every lane ends with the pairwise-tree product of the 64
// lane partials (strides 32, 16, ..., 1).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cplx64.hpp"
#include "kernels.hpp"

using namespace nq64;

namespace {

__device__ inline cd shfl_xor_cd(cd v, int m)
{
    return mk(__shfl_xor(v.re, m), __shfl_xor(v.im, m));
}

// all 64 lanes must be active
__device__ inline cd wave_prod(cd v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = mul(v, shfl_xor_cd(v, off));
    return v;
}

__device__ inline cd wave_sum(cd v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = add(v, shfl_xor_cd(v, off));
    return v;
}

__device__ inline int spin_bit(const uint32_t *sp, int NW, int k, int i)
{
    return (int)((sp[k * NW + (i >> 5)] >> (i & 31)) & 1u);
}

__global__ void k_tables64(int N, int M, int K, int L, const cd *__restrict__ W, cd *__restrict__ chT,
                           cd *__restrict__ shT, cd *__restrict__ eaP, cd *__restrict__ eaN)
{
    long long n = (long long)K * N * M;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x) {
        int j = (int)(idx % M);
        long long r = idx / M;
        int i = (int)(r % N);
        int kk = (int)(r / N);
        cd w2 = scale(2.0, W[(long long)kk * L + N + M + (long long)i * M + j]);
        chT[idx] = ccosh_(w2);
        shT[idx] = csinh_(w2);
        if (j == 0) {
            cd a2 = scale(2.0, W[(long long)kk * L + i]);
            eaP[kk * N + i] = cexp_(a2);
            eaN[kk * N + i] = cexp_(neg(a2));
        }
    }
}

// Chain start (rbm.cc:604-608): pcg64(seed + c), K*N Bernoulli draws
// (rbm.cc:99-119), state_fcn (rbm.cc:148-173) and the inverse.
// state_fcn (rbm.cc:148-173) of the spins in sp (LDS): tanh(theta) into tl, Phi and
// its inverse
template <int K, int JPL>
__device__ inline void state64(const Fp64SampleArgs &a, cd *tl, const uint32_t *sp, cd *phi, cd *phi_inv)
{
    const int N = a.N, M = a.M, NW = a.NW, L = a.L;
    const int lane = threadIdx.x;
    const cd *W = static_cast<const cd *>(a.W);
    {
        cd part[K * K];
#pragma unroll
        for (int q = 0; q < K * K; q++) part[q] = one();
#pragma unroll
        for (int q = 0; q < JPL; q++) {
            int j = lane + 64 * q;
            if (j < M) {
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    cd th[K];
                    cd bj = W[(long long)kk * L + N + j];
#pragma unroll
                    for (int k = 0; k < K; k++) th[k] = bj;
                    for (int i = 0; i < N; i++) {
                        cd w = W[(long long)kk * L + N + M + (long long)i * M + j];
#pragma unroll
                        for (int k = 0; k < K; k++) th[k] = add(th[k], spin_bit(sp, NW, k, i) ? neg(w) : w);
                    }
#pragma unroll
                    for (int k = 0; k < K; k++) {
                        tl[(k * K + kk) * M + j] = ctanh_(th[k]);
                        cd ch = ccosh_(th[k]);
                        part[k * K + kk] = (q == 0) ? ch : mul(part[k * K + kk], ch);
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++)
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                cd A = zero();
                for (int i = 0; i < N; i++) {
                    cd ai = W[(long long)kk * L + i];
                    A = add(A, spin_bit(sp, NW, k, i) ? neg(ai) : ai);
                }
                phi[k * K + kk] = mul(cexp_(A), wave_prod(part[k * K + kk]));
            }
        invK<K>(phi, phi_inv);
    }
    __syncthreads();
}

template <int K, int JPL>
__device__ inline void chain_start64(const Fp64SampleArgs &a, int c, cd *tl, uint32_t *sp, pcg &rng, cd *phi,
                                     cd *phi_inv)
{
    const int N = a.N, NW = a.NW;
    const int lane = threadIdx.x;
    rng = pcg_seed(a.seed + (uint64_t)c);
    for (int w = 0; w < K * NW; w++) {
        uint32_t word = 0;
        int k = w / NW, i0 = (w % NW) * 32;
        for (int b = 0; b < 32 && i0 + b < N; b++) {
            (void)k;
            if (canonical(pcg_next(rng)) < 0.5) word |= 1u << b;
        }
        if (lane == 0) sp[w] = word;
    }
    __syncthreads();
    state64<K, JPL>(a, tl, sp, phi, phi_inv);
}

// Width of an accept decision's near-tie window, relative to the acceptance
// probability min(1, p): 2^-40 (about 4096 ulps).  The amplitude ratio p is a product
// over M hidden units and a K x K determinant, so two implementations that round in a
// different order (the reference's libm, Eigen and FMA contraction against the
// canonical order here) can differ by many ulps of p; a uniform draw closer to
// min(1, p) than this window may be decided the other way by them.
#define NQ_TIE_REL 0x1p-40

// One Metropolis sweep (rbm.cc:530-552): proposals k < K, i < N in order.  ties
// counts the decisions whose draw falls within the near-tie window of min(1, p).
template <int K, int JPL>
__device__ inline void sweep64(const Fp64SampleArgs &a, cd *tl, uint32_t *sp, pcg &rng, cd *phi, cd *phi_inv,
                               unsigned long long &accepts, unsigned long long &ties)
{
    const int N = a.N, M = a.M, NW = a.NW;
    const int lane = threadIdx.x;
    const cd *chT = static_cast<const cd *>(a.chT);
    const cd *shT = static_cast<const cd *>(a.shT);
    const cd *eaP = static_cast<const cd *>(a.eaP);
    const cd *eaN = static_cast<const cd *>(a.eaN);
    cd det_phi = detK<K>(phi);
#pragma unroll
    for (int k = 0; k < K; k++) {
        for (int i = 0; i < N; i++) {
            const int si = spin_bit(sp, NW, k, i);
            cd tnew[JPL][K];
            cd part[K];
#pragma unroll
            for (int kk = 0; kk < K; kk++) part[kk] = one();
#pragma unroll
            for (int q = 0; q < JPL; q++) {
                int j = lane + 64 * q;
                if (j < M) {
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        long long ti = ((long long)kk * N + i) * M + j;
                        cd chv = chT[ti];
                        cd shv = shT[ti];
                        if (si) shv = neg(shv);
                        cd th = tl[(k * K + kk) * M + j];
                        cd denom = sub(chv, mul(th, shv));
                        tnew[q][kk] = div(sub(mul(th, chv), shv), denom);
                        part[kk] = (q == 0) ? denom : mul(part[kk], denom);
                    }
                }
            }
            cd prow[K];
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                cd e = si ? eaP[kk * N + i] : eaN[kk * N + i];
                cd T = wave_prod(part[kk]);
                prow[kk] = mul(mul(phi[k * K + kk], e), T);
            }
            cd phip[K * K];
#pragma unroll
            for (int r = 0; r < K; r++)
#pragma unroll
                for (int cc = 0; cc < K; cc++) phip[r * K + cc] = (r == k) ? prow[cc] : phi[r * K + cc];
            cd det_p = detK<K>(phip);
            double prob = (det_phi.re == 0.0 && det_phi.im == 0.0) ? 1.0 : norm(det_p) / norm(det_phi);
            double u = canonical(pcg_next(rng));
            const double pmin = prob < 1.0 ? prob : 1.0;
            if (fabs(u - pmin) <= pmin * NQ_TIE_REL) ties++;
            if (u <= pmin) {
                det_phi = det_p;
#pragma unroll
                for (int q = 0; q < K * K; q++) phi[q] = phip[q];
                invK<K>(phi, phi_inv);
#pragma unroll
                for (int q = 0; q < JPL; q++) {
                    int j = lane + 64 * q;
                    if (j < M)
#pragma unroll
                        for (int kk = 0; kk < K; kk++) tl[(k * K + kk) * M + j] = tnew[q][kk];
                }
                if (lane == 0) sp[k * NW + (i >> 5)] ^= 1u << (i & 31);
                accepts++;
                __syncthreads();
            }
        }
    }
}

// hs-like local energy (rbm.cc:1092-1119): res row k = sum over pairs i > j (i, then j
// ascending) of J_ij (2 Phi^(ij)_k - delta Phi_k) if s^k_i != s^k_j, else
// (J_ij delta) Phi_k; H = Phi^+ res.  Phi^(ij)_k: spin i, then spin j flipped in
// replica k (two flip_state_fcn steps, rbm.cc:1106-1108), the second on the angles
// left by the first; each product over hidden units is a lane product and a
// butterfly, as in the sweep.  Pairs with J_ij == 0 add exact zeros and are skipped.
template <int K, int JPL>
__device__ inline void hres_hs64(const Fp64SampleArgs &a, const cd *tl, const uint32_t *sp, const cd *phi, cd *res)
{
    const int N = a.N, M = a.M, NW = a.NW;
    const int lane = threadIdx.x;
    const cd *chT = static_cast<const cd *>(a.chT);
    const cd *shT = static_cast<const cd *>(a.shT);
    const cd *eaP = static_cast<const cd *>(a.eaP);
    const cd *eaN = static_cast<const cd *>(a.eaN);
    const double delta = a.delta;
#pragma unroll
    for (int q = 0; q < K * K; q++) res[q] = zero();
    for (int i = 1; i < N; i++)
        for (int j = 0; j < i; j++) {
            const double pij = a.J[(long long)i * N + j];
            if (pij == 0.0) continue;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int si = spin_bit(sp, NW, k, i), sj = spin_bit(sp, NW, k, j);
                if (si == sj) {
#pragma unroll
                    for (int kk = 0; kk < K; kk++) res[k * K + kk] = add(res[k * K + kk], scale(pij * delta, phi[k * K + kk]));
                    continue;
                }
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    cd t1[JPL];
                    cd part = one();
#pragma unroll
                    for (int q = 0; q < JPL; q++) {
                        const int hj = lane + 64 * q;
                        if (hj < M) {
                            const long long ti = ((long long)kk * N + i) * M + hj;
                            const cd chv = chT[ti];
                            cd shv = shT[ti];
                            if (si) shv = neg(shv);
                            const cd th = tl[(k * K + kk) * M + hj];
                            const cd den = sub(chv, mul(th, shv));
                            t1[q] = div(sub(mul(th, chv), shv), den);
                            part = (q == 0) ? den : mul(part, den);
                        }
                    }
                    const cd e = si ? eaP[kk * N + i] : eaN[kk * N + i];
                    cd v = mul(mul(phi[k * K + kk], e), wave_prod(part));
                    cd part2 = one();
#pragma unroll
                    for (int q = 0; q < JPL; q++) {
                        const int hj = lane + 64 * q;
                        if (hj < M) {
                            const long long tj = ((long long)kk * N + j) * M + hj;
                            cd shv = shT[tj];
                            if (sj) shv = neg(shv);
                            const cd den = sub(chT[tj], mul(t1[q], shv));
                            part2 = (q == 0) ? den : mul(part2, den);
                        }
                    }
                    const cd e2 = sj ? eaP[kk * N + j] : eaN[kk * N + j];
                    v = mul(mul(v, e2), wave_prod(part2));
                    const cd d = sub(scale(2.0, v), scale(delta, phi[k * K + kk]));
                    res[k * K + kk] = add(res[k * K + kk], scale(pij, d));
                }
            }
        }
}

// ising-like numerator (rbm.cc:1042-1064): res = diag(E_zz) Phi - h sum_i Phi^(i), E_zz
// of replica (lane % K) (rbm.cc:1012-1031); all 64 lanes active
template <int K, int JPL>
__device__ inline void hres_ising64(const Fp64SampleArgs &a, const cd *tl, const uint32_t *sp, const cd *phi,
                                    cd *res)
{
    const int N = a.N, M = a.M, NW = a.NW;
    const int lane = threadIdx.x;
    const cd *chT = static_cast<const cd *>(a.chT);
    const cd *shT = static_cast<const cd *>(a.shT);
    const cd *eaP = static_cast<const cd *>(a.eaP);
    const cd *eaN = static_cast<const cd *>(a.eaN);
    double ez;
    {
        int kr = lane % K;
        double r = 0.0;
        for (int i = 1; i < N; i++) {
            int bi = spin_bit(sp, NW, kr, i);
            for (int j = 0; j < i; j++) {
                double v = a.J[(long long)i * N + j];
                r += (bi ^ spin_bit(sp, NW, kr, j)) ? -v : v;
            }
        }
        ez = r * 2.0;
    }
    cd sx[K * K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double ek = __shfl(ez, k);
#pragma unroll
        for (int kk = 0; kk < K; kk++) res[k * K + kk] = scale(ek, phi[k * K + kk]);
    }
#pragma unroll
    for (int q = 0; q < K * K; q++) sx[q] = zero();
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int si = spin_bit(sp, NW, k, i);
#pragma unroll
            for (int kk = 0; kk < K; kk++) {
                cd part = one();
#pragma unroll
                for (int q = 0; q < JPL; q++) {
                    int j = lane + 64 * q;
                    if (j < M) {
                        long long ti = ((long long)kk * N + i) * M + j;
                        cd shv = shT[ti];
                        if (si) shv = neg(shv);
                        cd d = sub(chT[ti], mul(tl[(k * K + kk) * M + j], shv));
                        part = (q == 0) ? d : mul(part, d);
                    }
                }
                cd e = si ? eaP[kk * N + i] : eaN[kk * N + i];
                cd v = mul(mul(phi[k * K + kk], e), wave_prod(part));
                sx[k * K + kk] = add(sx[k * K + kk], v);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < K * K; q++) res[q] = sub(res[q], scale(a.h, sx[q]));
}

template <int K, int JPL>
__global__ __launch_bounds__(64) void k_sample64(Fp64SampleArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, M = a.M, NW = a.NW, L = a.L;
    cd *tl = reinterpret_cast<cd *>(smem);                      // [k][kk][j]
    uint32_t *sp = reinterpret_cast<uint32_t *>(tl + K * K * M); // [k][word]
    const int lane = threadIdx.x;
    const int c = blockIdx.x;
    cd *Hs = static_cast<cd *>(a.Hs);
    cd *Es = static_cast<cd *>(a.Es);

    cd phi[K * K], phi_inv[K * K];
    pcg rng;
    chain_start64<K, JPL>(a, c, tl, sp, rng, phi, phi_inv);
    unsigned long long accepts = 0, ties = 0;
    const int nsweeps = a.B + a.m;
    for (int sw = 0; sw < nsweeps; sw++) {
        sweep64<K, JPL>(a, tl, sp, rng, phi, phi_inv, accepts, ties);
        if (sw < a.B) continue;

        // ---- one sample
        const long long s = (long long)c * a.m + (sw - a.B);
        __syncthreads();

        // local energy: hs-like (rbm.cc:1092-1119) or ising-like (rbm.cc:1042-1064)
        cd H[K * K];
        {
            cd res[K * K];
            if (a.model != 0) hres_hs64<K, JPL>(a, tl, sp, phi, res);
            else hres_ising64<K, JPL>(a, tl, sp, phi, res);
            matmulK<K>(phi_inv, res, H);
        }
        cd energy = zero();
#pragma unroll
        for (int k = 0; k < K; k++) energy = add(energy, H[k * K + k]);
#pragma unroll
        for (int q = 0; q < K * K; q++)
            if (lane == q) Hs[s * K * K + q] = H[q];
        if (lane == 0) Es[s] = energy;
        for (int w = lane; w < K * NW; w += 64) a.spins[s * K * NW + w] = sp[w];

        // log-derivatives (rbm.cc:199-239).  Columns c != k of Phi^+ derPhi equal
        // those of G = Phi^+ Phi (same operands, same order), so only column k is
        // formed per entry.
        if (a.O) {
            cd G[K * K];
            matmulK<K>(phi_inv, phi, G);
            cd *Orow = static_cast<cd *>(a.O) + s * a.P;
            double *Orow_r = static_cast<double *>(a.O) + s * a.Pld;  // real weights: Re O only
            if (a.Oreal)
                for (long long p = a.P + lane; p < a.Pld; p += 64) Orow_r[p] = 0.0;
            for (long long p = lane; p < a.P; p += 64) {
                int k = (int)(p / L);
                int r = (int)(p % L);
                cd col[K];
                if (r < N) {
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        cd v = phi[kk * K + k];
                        col[kk] = spin_bit(sp, NW, kk, r) ? neg(v) : v;
                    }
                } else if (r < N + M) {
                    int j = r - N;
#pragma unroll
                    for (int kk = 0; kk < K; kk++) col[kk] = mul(phi[kk * K + k], tl[(kk * K + k) * M + j]);
                } else {
                    int rr = r - N - M;
                    int i = rr / M, j = rr % M;
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        cd v = mul(phi[kk * K + k], tl[(kk * K + k) * M + j]);
                        col[kk] = spin_bit(sp, NW, kk, i) ? neg(v) : v;
                    }
                }
                cd X[K * K];
#pragma unroll
                for (int rw = 0; rw < K; rw++) {
                    cd y = mul(phi_inv[rw * K], col[0]);
#pragma unroll
                    for (int l = 1; l < K; l++) y = add(y, mul(phi_inv[rw * K + l], col[l]));
#pragma unroll
                    for (int cc = 0; cc < K; cc++) X[rw * K + cc] = (cc == k) ? y : G[rw * K + cc];
                }
                const cd op = detK<K>(X);
                if (a.Oreal) Orow_r[p] = op.re;
                else Orow[p] = op;
            }
        }
        __syncthreads();
    }
    if (lane == 0) a.accepts[c] = accepts;
    if (lane == 0 && a.ties) a.ties[c] = ties;
}

// Dense S of the rbm-N* executables (rbm.cc:651-657, 686-696): oo[l][l'] =
// (1/S) sum_s conj(O_sl) O_sl' - conj(o_l) o_l'.  One 16 x 16 tile per block, samples
// in chunks of 16 staged in LDS, each element summed in ascending sample order.
__global__ __launch_bounds__(256) void k_dense_gram64(const cd *__restrict__ O, long long S, long long P,
                                                      const cd *__restrict__ o, cd *__restrict__ oo)
{
    __shared__ cd A[16][17], B[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long long l0 = (long long)blockIdx.y * 16, m0 = (long long)blockIdx.x * 16;
    cd acc = zero();
    for (long long s0 = 0; s0 < S; s0 += 16) {
        const long long s = s0 + ty;
        A[ty][tx] = (s < S && l0 + tx < P) ? O[s * P + l0 + tx] : zero();
        B[ty][tx] = (s < S && m0 + tx < P) ? O[s * P + m0 + tx] : zero();
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) acc = add(acc, mul(conj(A[q][ty]), B[q][tx]));
        __syncthreads();
    }
    const long long l = l0 + ty, m = m0 + tx;
    if (l < P && m < P) oo[l * P + m] = sub(divr(acc, (double)S), mul(conj(o[l]), o[m]));
}

// Dense operator (rbm.cc:686-696): out[l] = skk[l] v[l] + sum_l' oo[l][l'] v[l'].  One
// wavefront per row, lanes over l', butterfly sum.
__global__ __launch_bounds__(256) void k_dense_apply64(const cd *__restrict__ oo, const cd *__restrict__ skk,
                                                       const cd *__restrict__ v, long long P, cd *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const long long l = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (l >= P) return;
    cd acc = zero();
    for (long long q = lane; q < P; q += 64) acc = add(acc, mul(oo[l * P + q], v[q]));
    acc = wave_sum(acc);
    if (lane == 0) out[l] = add(mul(skk[l], v[l]), acc);
}

// Exhaustive evaluation for the --dump-states/ham/deriv files (rbm.cc:474-526,
// 1149-1172): one wavefront per multi-spin configuration cfg = cfg0 + blockIdx.x
// (replica k's spins are bits [k N, k N + N) of cfg, N K < 32).  Writes Phi (K*K),
// the local-energy numerator res (H_loc = Phi^+ res, K*K) and tanh(theta)
// ([k][kk][j], K*K*M); any output may be NULL.
template <int K, int JPL>
__global__ __launch_bounds__(64) void k_enum64(Fp64SampleArgs a, long long cfg0, cd *phi_out, cd *res_out,
                                               cd *tanh_out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, M = a.M;
    cd *tl = reinterpret_cast<cd *>(smem);
    uint32_t *sp = reinterpret_cast<uint32_t *>(tl + K * K * M);
    const int lane = threadIdx.x;
    const long long cfg = cfg0 + blockIdx.x;
    if (lane < K) sp[lane] = (uint32_t)((cfg >> (lane * N)) & ((1ll << N) - 1));
    __syncthreads();
    cd phi[K * K], phi_inv[K * K];
    state64<K, JPL>(a, tl, sp, phi, phi_inv);
    cd res[K * K];
    if (a.model != 0) hres_hs64<K, JPL>(a, tl, sp, phi, res);
    else hres_ising64<K, JPL>(a, tl, sp, phi, res);
    const long long e = (long long)blockIdx.x;
#pragma unroll
    for (int q = 0; q < K * K; q++)
        if (lane == q) {
            if (phi_out) phi_out[e * K * K + q] = phi[q];
            if (res_out) res_out[e * K * K + q] = res[q];
        }
    if (tanh_out)
        for (int q = lane; q < K * K * M; q += 64) tanh_out[e * K * K * M + q] = tl[q];
}

// accumulate the K x K matrix Phi^+ X into a chain's correlation slot (lane q
// owns entry q)
template <int K>
__device__ inline void cor_add(cd *slot, const cd *phi_inv, const cd *X)
{
    cd mat[K * K];
    matmulK<K>(phi_inv, X, mat);
    const int lane = threadIdx.x;
#pragma unroll
    for (int q = 0; q < K * K; q++)
        if (lane == q) slot[q] = add(slot[q], mat[q]);
}

// Correlation estimators (rbm.cc:730-771 with the operators of rbm.cc:774-829): per
// sample and index, Phi^+ fcn(index) accumulated per chain into
// cor[c][idx][K*K]: idx = i (sigma^x_i: spin i flipped in every replica), N + i
// (sigma^z_i: rows with spin i down negated), 2N + i*N + j (sigma^x_i sigma^x_j),
// 2N + N*N + i*N + j (sigma^z_i sigma^z_j).  Flipped amplitudes use the incremental
// update of flip_state_fcn applied once (x) or twice (xx) instead of state_fcn from
// scratch (same values, cheaper).
template <int K, int JPL>
__global__ __launch_bounds__(64) void k_corr64(Fp64SampleArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, M = a.M, NW = a.NW;
    cd *tl = reinterpret_cast<cd *>(smem);
    uint32_t *sp = reinterpret_cast<uint32_t *>(tl + K * K * M);
    const int lane = threadIdx.x;
    const int c = blockIdx.x;
    const cd *chT = static_cast<const cd *>(a.chT);
    const cd *shT = static_cast<const cd *>(a.shT);
    const cd *eaP = static_cast<const cd *>(a.eaP);
    const cd *eaN = static_cast<const cd *>(a.eaN);
    const long long ncor = 2LL * N + 2LL * N * N;
    cd *cor = static_cast<cd *>(a.cor) + (long long)c * ncor * K * K;
    for (long long q = lane; q < ncor * K * K; q += 64) cor[q] = zero();

    cd phi[K * K], phi_inv[K * K];
    pcg rng;
    chain_start64<K, JPL>(a, c, tl, sp, rng, phi, phi_inv);
    unsigned long long accepts = 0, ties = 0;
    for (int sw = 0; sw < a.B + a.m; sw++) {
        sweep64<K, JPL>(a, tl, sp, rng, phi, phi_inv, accepts, ties);
        if (sw < a.B) continue;
        __syncthreads();
        // sigma^x_i and sigma^z_i
        for (int i = 0; i < N; i++) {
            cd X[K * K], Z[K * K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int si = spin_bit(sp, NW, k, i);
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    cd part = one();
#pragma unroll
                    for (int q = 0; q < JPL; q++) {
                        int j = lane + 64 * q;
                        if (j < M) {
                            long long ti = ((long long)kk * N + i) * M + j;
                            cd shv = shT[ti];
                            if (si) shv = neg(shv);
                            cd d = sub(chT[ti], mul(tl[(k * K + kk) * M + j], shv));
                            part = (q == 0) ? d : mul(part, d);
                        }
                    }
                    cd e = si ? eaP[kk * N + i] : eaN[kk * N + i];
                    X[k * K + kk] = mul(mul(phi[k * K + kk], e), wave_prod(part));
                    Z[k * K + kk] = si ? neg(phi[k * K + kk]) : phi[k * K + kk];
                }
            }
            cor_add<K>(cor + (long long)i * K * K, phi_inv, X);
            cor_add<K>(cor + (long long)(N + i) * K * K, phi_inv, Z);
        }
        // sigma^x_i sigma^x_j and sigma^z_i sigma^z_j
        for (int i = 0; i < N; i++) {
            for (int jj = 0; jj < N; jj++) {
                cd Z[K * K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int f = spin_bit(sp, NW, k, i) ^ spin_bit(sp, NW, k, jj);
#pragma unroll
                    for (int kk = 0; kk < K; kk++) Z[k * K + kk] = f ? neg(phi[k * K + kk]) : phi[k * K + kk];
                }
                cor_add<K>(cor + (2LL * N + (long long)N * N + (long long)i * N + jj) * K * K, phi_inv, Z);
                if (jj < i) continue;
                if (jj == i) {
                    cor_add<K>(cor + (2LL * N + (long long)i * N + i) * K * K, phi_inv, phi);
                    continue;
                }
                cd X[K * K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int si = spin_bit(sp, NW, k, i), sj = spin_bit(sp, NW, k, jj);
#pragma unroll
                    for (int kk = 0; kk < K; kk++) {
                        cd part = one();
#pragma unroll
                        for (int q = 0; q < JPL; q++) {
                            int j = lane + 64 * q;
                            if (j < M) {
                                long long ti = ((long long)kk * N + i) * M + j;
                                long long tj = ((long long)kk * N + jj) * M + j;
                                cd t = tl[(k * K + kk) * M + j];
                                cd shi = shT[ti], shj = shT[tj];
                                if (si) shi = neg(shi);
                                if (sj) shj = neg(shj);
                                cd d1 = sub(chT[ti], mul(t, shi));
                                cd t1 = div(sub(mul(t, chT[ti]), shi), d1);
                                cd d2 = sub(chT[tj], mul(t1, shj));
                                cd d = mul(d1, d2);
                                part = (q == 0) ? d : mul(part, d);
                            }
                        }
                        cd ei = si ? eaP[kk * N + i] : eaN[kk * N + i];
                        cd ej = sj ? eaP[kk * N + jj] : eaN[kk * N + jj];
                        X[k * K + kk] = mul(mul(mul(phi[k * K + kk], ei), ej), wave_prod(part));
                    }
                }
                cor_add<K>(cor + (2LL * N + (long long)i * N + jj) * K * K, phi_inv, X);
                cor_add<K>(cor + (2LL * N + (long long)jj * N + i) * K * K, phi_inv, X);
            }
        }
        __syncthreads();
    }
}

// Column sums over rows [s0, s1) of chunk blockIdx.y; thread = column.
__global__ void k_moments64(const cd *__restrict__ O, const cd *__restrict__ E, long long S, long long P,
                            int nchunk, cd *__restrict__ po, cd *__restrict__ ps, cd *__restrict__ pf)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    long long s0 = S * blockIdx.y / nchunk, s1 = S * (blockIdx.y + 1) / nchunk;
    cd so = zero(), ss = zero(), sf = zero();
    for (long long s = s0; s < s1; s++) {
        cd v = O[s * P + p];
        so = add(so, v);
        ss = add(ss, mul(v, conj(v)));
        sf = add(sf, mul(E[s], conj(v)));
    }
    long long o = (long long)blockIdx.y * P + p;
    po[o] = so;
    ps[o] = ss;
    pf[o] = sf;
}

// The same sums for real O (real weights: every imaginary part of O is zero), rows of
// stride ld doubles.  Each sum takes the real operations the complex one reduces to
// when Im O = 0, in the same order, so the partials equal k_moments64's bit for bit:
// so = (sum O, 0), ss = (sum O*O, 0), sf = (sum E.re*O, sum E.im*O).
__global__ void k_moments64r(const double *__restrict__ O, long long ld, const cd *__restrict__ E, long long S,
                             long long P, int nchunk, cd *__restrict__ po, cd *__restrict__ ps,
                             cd *__restrict__ pf)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    long long s0 = S * blockIdx.y / nchunk, s1 = S * (blockIdx.y + 1) / nchunk;
    double so = 0.0, ss = 0.0, fr = 0.0, fi = 0.0;
    for (long long s = s0; s < s1; s++) {
        const double v = O[s * ld + p];
        const cd e = E[s];
        so = so + v;
        ss = ss + v * v;
        fr = fr + e.re * v;
        fi = fi + e.im * v;
    }
    long long o = (long long)blockIdx.y * P + p;
    po[o] = mk(so, 0.0);
    ps[o] = mk(ss, 0.0);
    pf[o] = mk(fr, fi);
}

// out[col] = sum over chunks of part[chunk][col], chunks in order
__global__ void k_chunksum64(const cd *__restrict__ part, long long ncol, int nchunk, cd *__restrict__ out)
{
    long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    cd acc = part[col];
    for (int ch = 1; ch < nchunk; ch++) acc = add(acc, part[(long long)ch * ncol + col]);
    out[col] = acc;
}

// the same sum for few columns: block per column, thread t sums chunks t, t + 256, ...
// in ascending order, then a fixed pairwise tree over the 256 partials
__global__ __launch_bounds__(256) void k_chunksum_tree64(const cd *__restrict__ part, long long ncol, int nchunk,
                                                         cd *__restrict__ out)
{
    __shared__ cd red[256];
    const long long col = blockIdx.x;
    cd acc = zero();
    for (int ch = threadIdx.x; ch < nchunk; ch += 256) acc = add(acc, part[(long long)ch * ncol + col]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = add(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[col] = red[0];
}

__global__ void k_colsum64(const cd *__restrict__ X, long long S, long long ncol, int nchunk,
                           cd *__restrict__ part)
{
    long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    long long s0 = S * blockIdx.y / nchunk, s1 = S * (blockIdx.y + 1) / nchunk;
    cd acc = zero();
    // rows added in order; the loads run eight ahead of the adds
    long long s = s0;
    for (; s + 8 <= s1; s += 8) {
        cd x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = X[(s + u) * ncol + col];
#pragma unroll
        for (int u = 0; u < 8; u++) acc = add(acc, x[u]);
    }
    for (; s < s1; s++) acc = add(acc, X[s * ncol + col]);
    part[(long long)blockIdx.y * ncol + col] = acc;
}

// w_s = sum_p O_sp v_p; one wave per row, 4 waves per block, grid-stride over rows
__global__ __launch_bounds__(256) void k_sv_rowdot64(const cd *__restrict__ O, const cd *__restrict__ v,
                                                     long long S, long long P, cd *__restrict__ w)
{
    const int lane = threadIdx.x & 63;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    for (long long s = wave; s < S; s += nwaves) {
        const cd *row = O + s * P;
        cd acc = zero();
        for (long long p = lane; p < P; p += 64) acc = add(acc, mul(row[p], v[p]));
        acc = wave_sum(acc);
        if (lane == 0) w[s] = acc;
    }
}

// part[chunk][p] = sum_{s in chunk} w_s conj(O_sp)
__global__ void k_sv_colacc64(const cd *__restrict__ O, const cd *__restrict__ w, long long S, long long P,
                              int nchunk, cd *__restrict__ part)
{
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    long long s0 = S * blockIdx.y / nchunk, s1 = S * (blockIdx.y + 1) / nchunk;
    cd acc = zero();
    for (long long s = s0; s < s1; s++) acc = add(acc, mul(w[s], conj(O[s * P + p])));
    part[(long long)blockIdx.y * P + p] = acc;
}

// ---------------------------------------------------------------- S*v, one pass
// FP64 O rows as 16-byte vectors: 2 real entries (real weights: O is real) or 1
// complex entry.
template <bool CPLX> struct Row64;
template <> struct Row64<false> {
    using D = double;
    static constexpr int W = 2;
};
template <> struct Row64<true> {
    using D = cd;
    static constexpr int W = 1;
};

__device__ inline double d_add(double a, double b) { return a + b; }
__device__ inline cd d_add(cd a, cd b) { return add(a, b); }
__device__ inline double d_sub(double a, double b) { return a - b; }
__device__ inline cd d_sub(cd a, cd b) { return sub(a, b); }
__device__ inline double d_mul(double a, double b) { return a * b; }
__device__ inline cd d_mul(cd a, cd b) { return mul(a, b); }
__device__ inline double d_conj(double a) { return a; }
__device__ inline cd d_conj(cd a) { return conj(a); }
__device__ inline double d_xor(double a, int m) { return __shfl_xor(a, m); }
__device__ inline cd d_xor(cd a, int m) { return shfl_xor_cd(a, m); }
template <typename D> __device__ inline D d_zero();
template <> __device__ inline double d_zero<double>() { return 0.0; }
template <> __device__ inline cd d_zero<cd>() { return zero(); }

template <bool CPLX>
__device__ inline void row_vec(const double2 &x, typename Row64<CPLX>::D *o)
{
    if constexpr (CPLX) {
        o[0] = mk(x.x, x.y);
    } else {
        o[0] = x.x;
        o[1] = x.y;
    }
}

typedef double d2v __attribute__((ext_vector_type(2)));

// O is streamed once per S*v: loads that do not keep it in the caches
__device__ __forceinline__ double2 ld_stream2(const double2 *p)
{
    d2v r = __builtin_nontemporal_load((const d2v *)p);
    return make_double2(r.x, r.y);
}

// S*v in one streamed pass over FP64 O rows (rbm.cc:661-684).  A block of WPR
// wavefronts owns one row at a time (rows blockIdx.x, blockIdx.x + gridDim.x, ...);
// lane l of wave h holds the row's 16-byte vectors u = l + 64 (h + WPR q), q < QV.
// Per row: w_s = O_s^T v - o^T v (v staged once per block in LDS; per-lane partial
// dot in ascending q, 64-lane butterfly, the WPR wave partials added in wave order
// through LDS, one barrier per row on alternating slots), then acc += conj(O_s) w_s
// from the same registers, so O is read once.  The next row's loads are issued before
// the current row's arithmetic.  Every wave of a block takes the same rows, so the
// barriers match.  Output: the block's partial row (its own vectors, no cross-wave
// sum) and its sum of w_s.  Entries of v at p >= P and O's row padding are zero.
template <bool CPLX, int QV, int WPR>
__global__ __launch_bounds__(64 * WPR) void k_sv_rows64(const double2 *__restrict__ O, const void *v_,
                                                         const void *ov_, long long S, long long nvec, long long P,
                                                         void *part_, void *wsum_)
{
    using D = typename Row64<CPLX>::D;
    constexpr int W = Row64<CPLX>::W;
    extern __shared__ __align__(16) unsigned char smem_sv64[];
    __shared__ D red[2][WPR];
    D *vl = (D *)smem_sv64;
    const D *v = (const D *)v_;
    const D ov = *(const D *)ov_;
    const int t = threadIdx.x, lane = t & 63, h = t >> 6;
    for (long long q = t; q < nvec * W; q += blockDim.x) vl[q] = q < P ? v[q] : d_zero<D>();
    __syncthreads();
    D acc[QV][W];
#pragma unroll
    for (int q = 0; q < QV; q++)
#pragma unroll
        for (int e = 0; e < W; e++) acc[q][e] = d_zero<D>();
    D wacc = d_zero<D>();
    const double2 z2 = make_double2(0.0, 0.0);
    const long long ns = gridDim.x;
    long long s = blockIdx.x;
    double2 x[QV];
#pragma unroll
    for (int q = 0; q < QV; q++) {
        const long long u = lane + 64LL * (h + WPR * q);
        x[q] = (s < S && u < nvec) ? ld_stream2(O + s * nvec + u) : z2;
    }
    int par = 0;
    for (; s < S; s += ns) {
        double2 y[QV];
        const long long sn = s + ns;
#pragma unroll
        for (int q = 0; q < QV; q++) {
            const long long u = lane + 64LL * (h + WPR * q);
            y[q] = (sn < S && u < nvec) ? ld_stream2(O + sn * nvec + u) : z2;
        }
        D dot = d_zero<D>();
#pragma unroll
        for (int q = 0; q < QV; q++) {
            const long long u = lane + 64LL * (h + WPR * q);
            if (u < nvec) {
                D od[W];
                row_vec<CPLX>(x[q], od);
#pragma unroll
                for (int e = 0; e < W; e++) dot = d_add(dot, d_mul(od[e], vl[u * W + e]));
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) dot = d_add(dot, d_xor(dot, off));
        if constexpr (WPR > 1) {
            if (lane == 0) red[par][h] = dot;
            __syncthreads();
            dot = red[par][0];
#pragma unroll
            for (int k = 1; k < WPR; k++) dot = d_add(dot, red[par][k]);
            par ^= 1;
        }
        const D w = d_sub(dot, ov);
        wacc = d_add(wacc, w);
#pragma unroll
        for (int q = 0; q < QV; q++) {
            D od[W];
            row_vec<CPLX>(x[q], od);
#pragma unroll
            for (int e = 0; e < W; e++) acc[q][e] = d_add(acc[q][e], d_mul(w, d_conj(od[e])));
        }
#pragma unroll
        for (int q = 0; q < QV; q++) x[q] = y[q];
    }
    D *part = (D *)part_ + (long long)blockIdx.x * nvec * W;
#pragma unroll
    for (int q = 0; q < QV; q++) {
        const long long u = lane + 64LL * (h + WPR * q);
        if (u < nvec)
#pragma unroll
            for (int e = 0; e < W; e++) part[u * W + e] = acc[q][e];
    }
    if (t == 0) ((D *)wsum_)[blockIdx.x] = wacc;
}

// o^T v in one block: thread t sums p = t, t + 256, ... in ascending order, then a
// fixed pairwise tree over the 256 partials (the same value on every call)
template <bool CPLX>
__global__ __launch_bounds__(256) void k_dot64(const void *o_, const void *v_, long long P, void *out_)
{
    using D = typename Row64<CPLX>::D;
    __shared__ D red[256];
    const D *o = (const D *)o_, *v = (const D *)v_;
    D acc = d_zero<D>();
    for (long long p = threadIdx.x; p < P; p += 256) acc = d_add(acc, d_mul(o[p], v[p]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] = d_add(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *(D *)out_ = red[0];
}

// S v from the raw sums (rbm.cc:661-684, centred form):
// out[p] = (raw[p] - conj(o_p) raw[P]) / S + skk_p v_p
template <bool CPLX>
__global__ void k_sv_finish64(const void *raw_, const void *o_, const void *skk_, const void *v_, long long P,
                              double Sd, void *out_)
{
    using D = typename Row64<CPLX>::D;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const D *raw = (const D *)raw_, *o = (const D *)o_, *skk = (const D *)skk_, *v = (const D *)v_;
    D r = d_sub(raw[p], d_mul(d_conj(o[p]), raw[P]));
    if constexpr (CPLX) r = divr(r, Sd);
    else r = r / Sd;
    ((D *)out_)[p] = d_add(r, d_mul(skk[p], v[p]));
}

template <bool CPLX, int QV, int WPR>
hipError_t launch_rows64(const void *O, const void *v, const void *ov, long long S, long long nvec, long long P,
                         int nblk, void *part, void *wsum, hipStream_t st)
{
    const size_t lds = (size_t)nvec * 16;
    if (lds > 65536) {
        hipError_t e = nq_allow_dynamic_lds((const void *)k_sv_rows64<CPLX, QV, WPR>, lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_sv_rows64<CPLX, QV, WPR>), dim3(nblk), dim3(64 * WPR), lds, st, (const double2 *)O, v, ov,
                       S, nvec, P, part, wsum);
    return hipGetLastError();
}

template <bool CPLX>
hipError_t launch_rows64_shape(const Sv64Shape &sh, const void *O, const void *v, const void *ov, long long S,
                               long long nvec, long long P, void *part, void *wsum, hipStream_t st)
{
#define NQ_R64(q, w)                                                                                              \
    if (sh.qv == q && sh.wpr == w) return launch_rows64<CPLX, q, w>(O, v, ov, S, nvec, P, sh.blocks, part, wsum, st);
    NQ_R64(1, 1) NQ_R64(2, 1) NQ_R64(3, 1) NQ_R64(4, 1) NQ_R64(5, 1) NQ_R64(6, 1) NQ_R64(7, 1) NQ_R64(8, 1)
    NQ_R64(9, 1)
    NQ_R64(5, 2) NQ_R64(6, 2) NQ_R64(7, 2) NQ_R64(8, 2) NQ_R64(9, 2)
    NQ_R64(5, 4) NQ_R64(6, 4) NQ_R64(7, 4) NQ_R64(8, 4) NQ_R64(9, 4)
    NQ_R64(5, 8) NQ_R64(6, 8) NQ_R64(7, 8) NQ_R64(8, 8) NQ_R64(9, 8)
    NQ_R64(5, 16) NQ_R64(6, 16) NQ_R64(7, 16) NQ_R64(8, 16) NQ_R64(9, 16)
#undef NQ_R64
    return hipErrorInvalidValue;
}

template <int K, int JPL>
hipError_t launch_sample_t(const Fp64SampleArgs &a, hipStream_t st)
{
    size_t lds = sizeof(cd) * K * K * a.M + sizeof(uint32_t) * K * a.NW;
    hipLaunchKernelGGL((k_sample64<K, JPL>), dim3(a.C), dim3(64), lds, st, a);
    return hipGetLastError();
}

template <int K>
hipError_t launch_sample_k(const Fp64SampleArgs &a, hipStream_t st)
{
    int jpl = (a.M + 63) / 64;
    if (jpl <= 1) return launch_sample_t<K, 1>(a, st);
    if (jpl <= 2) return launch_sample_t<K, 2>(a, st);
    if (jpl <= 4) return launch_sample_t<K, 4>(a, st);
    if (jpl <= 8) return launch_sample_t<K, 8>(a, st);
    return hipErrorInvalidValue;
}


template <int K, int JPL>
hipError_t launch_enum_t(const Fp64SampleArgs &a, long long cfg0, long long ncfg, void *phi, void *res, void *tanh_,
                         hipStream_t st)
{
    size_t lds = sizeof(cd) * K * K * a.M + sizeof(uint32_t) * K * a.NW;
    hipLaunchKernelGGL((k_enum64<K, JPL>), dim3((unsigned)ncfg), dim3(64), lds, st, a, cfg0, (cd *)phi, (cd *)res,
                       (cd *)tanh_);
    return hipGetLastError();
}

template <int K>
hipError_t launch_enum_k(const Fp64SampleArgs &a, long long cfg0, long long ncfg, void *phi, void *res, void *tanh_,
                         hipStream_t st)
{
    int jpl = (a.M + 63) / 64;
    if (jpl <= 1) return launch_enum_t<K, 1>(a, cfg0, ncfg, phi, res, tanh_, st);
    if (jpl <= 2) return launch_enum_t<K, 2>(a, cfg0, ncfg, phi, res, tanh_, st);
    if (jpl <= 4) return launch_enum_t<K, 4>(a, cfg0, ncfg, phi, res, tanh_, st);
    if (jpl <= 8) return launch_enum_t<K, 8>(a, cfg0, ncfg, phi, res, tanh_, st);
    return hipErrorInvalidValue;
}

template <int K, int JPL>
hipError_t launch_corr_t(const Fp64SampleArgs &a, hipStream_t st)
{
    size_t lds = sizeof(cd) * K * K * a.M + sizeof(uint32_t) * K * a.NW;
    hipLaunchKernelGGL((k_corr64<K, JPL>), dim3(a.C), dim3(64), lds, st, a);
    return hipGetLastError();
}

template <int K>
hipError_t launch_corr_k(const Fp64SampleArgs &a, hipStream_t st)
{
    int jpl = (a.M + 63) / 64;
    if (jpl <= 1) return launch_corr_t<K, 1>(a, st);
    if (jpl <= 2) return launch_corr_t<K, 2>(a, st);
    if (jpl <= 4) return launch_corr_t<K, 4>(a, st);
    if (jpl <= 8) return launch_corr_t<K, 8>(a, st);
    return hipErrorInvalidValue;
}

} // namespace

hipError_t nq_launch_corr64(int K, const Fp64SampleArgs &a, hipStream_t st)
{
    switch (K) {
    case 1: return launch_corr_k<1>(a, st);
    case 2: return launch_corr_k<2>(a, st);
    case 3: return launch_corr_k<3>(a, st);
    case 4: return launch_corr_k<4>(a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t nq_launch_tables64(int N, int M, int K, int L, const void *W, void *chT, void *shT, void *eaP,
                              void *eaN, hipStream_t st)
{
    long long n = (long long)K * N * M;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_tables64, dim3(blocks), dim3(256), 0, st, N, M, K, L, (const cd *)W, (cd *)chT,
                       (cd *)shT, (cd *)eaP, (cd *)eaN);
    return hipGetLastError();
}

hipError_t nq_launch_sample64(int K, const Fp64SampleArgs &a, hipStream_t st)
{
    switch (K) {
    case 1: return launch_sample_k<1>(a, st);
    case 2: return launch_sample_k<2>(a, st);
    case 3: return launch_sample_k<3>(a, st);
    case 4: return launch_sample_k<4>(a, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t nq_launch_enum64(int K, const Fp64SampleArgs &a, long long cfg0, long long ncfg, void *phi, void *res,
                           void *tanh_, hipStream_t st)
{
    switch (K) {
    case 1: return launch_enum_k<1>(a, cfg0, ncfg, phi, res, tanh_, st);
    case 2: return launch_enum_k<2>(a, cfg0, ncfg, phi, res, tanh_, st);
    case 3: return launch_enum_k<3>(a, cfg0, ncfg, phi, res, tanh_, st);
    case 4: return launch_enum_k<4>(a, cfg0, ncfg, phi, res, tanh_, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t nq_launch_dense_gram64(const void *O, long long S, long long P, const void *o, void *oo, hipStream_t st)
{
    const unsigned nt = (unsigned)((P + 15) / 16);
    hipLaunchKernelGGL(k_dense_gram64, dim3(nt, nt), dim3(256), 0, st, (const cd *)O, S, P, (const cd *)o, (cd *)oo);
    return hipGetLastError();
}

hipError_t nq_launch_dense_apply64(const void *oo, const void *skk, const void *v, long long P, void *out,
                                   hipStream_t st)
{
    hipLaunchKernelGGL(k_dense_apply64, dim3((unsigned)((P + 3) / 4)), dim3(256), 0, st, (const cd *)oo,
                       (const cd *)skk, (const cd *)v, P, (cd *)out);
    return hipGetLastError();
}

hipError_t nq_launch_moments64(const void *O, const void *E, long long S, long long P, int nchunk, void *po,
                               void *ps, void *pf, void *out_o, void *out_s, void *out_f, hipStream_t st)
{
    dim3 g((unsigned)((P + 255) / 256), (unsigned)nchunk);
    hipLaunchKernelGGL(k_moments64, g, dim3(256), 0, st, (const cd *)O, (const cd *)E, S, P, nchunk, (cd *)po,
                       (cd *)ps, (cd *)pf);
    dim3 g2((unsigned)((P + 255) / 256));
    hipLaunchKernelGGL(k_chunksum64, g2, dim3(256), 0, st, (const cd *)po, P, nchunk, (cd *)out_o);
    hipLaunchKernelGGL(k_chunksum64, g2, dim3(256), 0, st, (const cd *)ps, P, nchunk, (cd *)out_s);
    hipLaunchKernelGGL(k_chunksum64, g2, dim3(256), 0, st, (const cd *)pf, P, nchunk, (cd *)out_f);
    return hipGetLastError();
}

hipError_t nq_launch_colsum64(const void *X, long long S, long long ncol, int nchunk, void *part, void *out,
                              hipStream_t st)
{
    dim3 g((unsigned)((ncol + 63) / 64), (unsigned)nchunk);
    hipLaunchKernelGGL(k_colsum64, g, dim3(64), 0, st, (const cd *)X, S, ncol, nchunk, (cd *)part);
    if (ncol <= 64)
        hipLaunchKernelGGL(k_chunksum_tree64, dim3((unsigned)ncol), dim3(256), 0, st, (const cd *)part, ncol, nchunk,
                           (cd *)out);
    else
        hipLaunchKernelGGL(k_chunksum64, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, st, (const cd *)part, ncol,
                           nchunk, (cd *)out);
    return hipGetLastError();
}

Sv64Shape nq_sv64_shape(long long nvec, long long S, size_t lds_max)
{
    Sv64Shape sh{0, 0, 0, 0};
    if (nvec < 1 || (size_t)nvec * 16 > lds_max) return sh;
    for (int wpr = 1; wpr <= 16; wpr *= 2) {
        const long long q = (nvec + 64LL * wpr - 1) / (64LL * wpr);
        if (q <= 9) {
            sh.wpr = wpr;
            sh.qv = (int)q;
            break;
        }
    }
    if (!sh.qv) return sh;
    if (sh.wpr > 1 && sh.qv < 5) sh.qv = 5;  // instantiated shapes only; idle slots are masked
    long long nb = 2048 / sh.wpr;
    if (nb > S) nb = S > 0 ? S : 1;
    sh.blocks = (int)nb;
    sh.ok = 1;
    return sh;
}

hipError_t nq_launch_sv_rows64(bool cplx, const Sv64Shape &sh, const void *O, const void *v, const void *ov,
                               long long S, long long nvec, long long P, void *part, void *wsum, void *scratch,
                               void *out, hipStream_t st)
{
    if (!sh.ok) return hipErrorInvalidValue;
    hipError_t e = cplx ? launch_rows64_shape<true>(sh, O, v, ov, S, nvec, P, part, wsum, st)
                        : launch_rows64_shape<false>(sh, O, v, ov, S, nvec, P, part, wsum, st);
    if (e != hipSuccess) return e;
    const size_t dsz = cplx ? 16 : 8;
    const long long stride = nvec * (cplx ? 1 : 2);
    e = nq_chunk_reduce(cplx, part, P, stride, sh.blocks, scratch, out, st);
    if (e != hipSuccess) return e;
    char *scr2 = static_cast<char *>(scratch) + (size_t)(sh.blocks / 32 + 1) * P * dsz;
    return nq_chunk_reduce(cplx, wsum, 1, 1, sh.blocks, scr2, static_cast<char *>(out) + (size_t)P * dsz, st);
}

hipError_t nq_launch_dot64(bool cplx, const void *o, const void *v, long long P, void *out, hipStream_t st)
{
    if (cplx) hipLaunchKernelGGL(k_dot64<true>, dim3(1), dim3(256), 0, st, o, v, P, out);
    else hipLaunchKernelGGL(k_dot64<false>, dim3(1), dim3(256), 0, st, o, v, P, out);
    return hipGetLastError();
}

hipError_t nq_launch_sv_finish64(bool cplx, const void *raw, const void *o, const void *skk, const void *v,
                                 long long P, double S, void *out, hipStream_t st)
{
    const dim3 g((unsigned)((P + 255) / 256));
    if (cplx) hipLaunchKernelGGL(k_sv_finish64<true>, g, dim3(256), 0, st, raw, o, skk, v, P, S, out);
    else hipLaunchKernelGGL(k_sv_finish64<false>, g, dim3(256), 0, st, raw, o, skk, v, P, S, out);
    return hipGetLastError();
}

hipError_t nq_launch_moments64r(const void *O, long long ld, const void *E, long long S, long long P, int nchunk,
                                void *po, void *ps, void *pf, void *out_o, void *out_s, void *out_f, hipStream_t st)
{
    dim3 g((unsigned)((P + 255) / 256), (unsigned)nchunk);
    hipLaunchKernelGGL(k_moments64r, g, dim3(256), 0, st, (const double *)O, ld, (const cd *)E, S, P, nchunk,
                       (cd *)po, (cd *)ps, (cd *)pf);
    dim3 g2((unsigned)((P + 255) / 256));
    hipLaunchKernelGGL(k_chunksum64, g2, dim3(256), 0, st, (const cd *)po, P, nchunk, (cd *)out_o);
    hipLaunchKernelGGL(k_chunksum64, g2, dim3(256), 0, st, (const cd *)ps, P, nchunk, (cd *)out_s);
    hipLaunchKernelGGL(k_chunksum64, g2, dim3(256), 0, st, (const cd *)pf, P, nchunk, (cd *)out_f);
    return hipGetLastError();
}

hipError_t nq_launch_sv64(const void *O, const void *v, long long S, long long P, int nchunk, void *w,
                          void *part, void *out, hipStream_t st)
{
    long long blocks = (S + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_sv_rowdot64, dim3((unsigned)blocks), dim3(256), 0, st, (const cd *)O, (const cd *)v, S,
                       P, (cd *)w);
    dim3 g((unsigned)((P + 255) / 256), (unsigned)nchunk);
    hipLaunchKernelGGL(k_sv_colacc64, g, dim3(256), 0, st, (const cd *)O, (const cd *)w, S, P, nchunk,
                       (cd *)part);
    hipLaunchKernelGGL(k_chunksum64, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, (const cd *)part, P,
                       nchunk, (cd *)out);
    return hipGetLastError();
}
