// kernels.hpp -- launch interface between the host library and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <shared_mutex>

// Stream capture and the calls that can end a capture in progress on another thread
// (device and pinned allocation, synchronous copies, stream and event creation, kernel
// attributes) exclude each other: a capture holds nq_capture_mutex() exclusively, such
// calls hold it shared through NqUnsafeHold.  A thread that is capturing (nq_capturing())
// does not take it again.
std::shared_mutex &nq_capture_mutex();
bool &nq_capturing();
class NqUnsafeHold {
  public:
    NqUnsafeHold()
    {
        if (!nq_capturing()) {
            nq_capture_mutex().lock_shared();
            held_ = true;
        }
    }
    ~NqUnsafeHold()
    {
        if (held_) nq_capture_mutex().unlock_shared();
    }
    NqUnsafeHold(const NqUnsafeHold &) = delete;
    NqUnsafeHold &operator=(const NqUnsafeHold &) = delete;

  private:
    bool held_ = false;
};

// Lets kernel `fn` launch with `bytes` of dynamic LDS on the current device.  The
// attribute is set once per kernel and device, to the most the kernel can take, never
// per launch: kernel attributes are shared by every thread of the process, so a
// per-launch value set by one thread could undercut another thread's launch in
// between.  hipErrorInvalidValue if `bytes` exceeds what the kernel can take.
hipError_t nq_allow_dynamic_lds(const void *fn, size_t bytes);

// Arguments of the FP64 sampler (one wavefront per chain).  Complex arrays are
// interleaved doubles.  Sample s = chain * m + t.
struct Fp64SampleArgs {
    int N, M, NW, L;
    long long P;
    int C, m, B;
    uint64_t seed;
    double h;
    const double *J;           // N x N
    int model;                 // 0 ising-like (rbm.cc:1042), 1 / 2 hs-like, delta = 1 / -1 (rbm.cc:1092)
    double delta;
    const void *W;             // P complex
    const void *chT, *shT;     // K*N*M complex: cosh(2w), sinh(2w), [kk][i][j]
    const void *eaP, *eaN;     // K*N complex: exp(2a), exp(-2a), [kk][i]
    void *O;                   // S rows of O (NULL: not stored): P complex, or Pld doubles when Oreal
    int Oreal;                 // 1: real weights, store Re O only (rows padded with zeros to Pld)
    long long Pld;
    void *Hs;                  // S*K*K complex
    void *Es;                  // S complex: trace H_loc
    uint32_t *spins;           // S*K*NW
    unsigned long long *accepts; // C
    unsigned long long *ties;  // C: accept decisions within the near-tie window (NULL: not written)
    void *cor;                 // correlation kernel: C*(2N+2N^2)*K*K complex
};

hipError_t nq_launch_tables64(int N, int M, int K, int L, const void *W, void *chT, void *shT, void *eaP,
                              void *eaN, hipStream_t st);
hipError_t nq_launch_sample64(int K, const Fp64SampleArgs &a, hipStream_t st);
// every multi-spin configuration [cfg0, cfg0 + ncfg) (N K < 32): Phi, the local-energy
// numerator res (H_loc = Phi^+ res) and tanh(theta); outputs may be NULL
hipError_t nq_launch_enum64(int K, const Fp64SampleArgs &a, long long cfg0, long long ncfg, void *phi, void *res,
                           void *tanh_, hipStream_t st);
// dense S (FP64): oo = <conj(O) O^T> - conj(o) o^T, and out = skk .* v + oo v
hipError_t nq_launch_dense_gram64(const void *O, long long S, long long P, const void *o, void *oo, hipStream_t st);
hipError_t nq_launch_dense_apply64(const void *oo, const void *skk, const void *v, long long P, void *out,
                                   hipStream_t st);
hipError_t nq_launch_corr64(int K, const Fp64SampleArgs &a, hipStream_t st);
hipError_t nq_launch_moments64(const void *O, const void *E, long long S, long long P, int nchunk, void *po,
                               void *ps, void *pf, void *out_o, void *out_s, void *out_f, hipStream_t st);
hipError_t nq_launch_colsum64(const void *X, long long S, long long ncol, int nchunk, void *part, void *out,
                              hipStream_t st);
hipError_t nq_launch_sv64(const void *O, const void *v, long long S, long long P, int nchunk, void *w,
                          void *part, void *out, hipStream_t st);
// moments of real O (rows of stride ld doubles), partials equal to nq_launch_moments64's
hipError_t nq_launch_moments64r(const void *O, long long ld, const void *E, long long S, long long P, int nchunk,
                                void *po, void *ps, void *pf, void *out_o, void *out_s, void *out_f, hipStream_t st);
// One-pass FP64 S*v over rows of nvec 16-byte vectors (2 real or 1 complex entries):
// wpr waves per row, qv vectors per lane, `blocks` partial rows.  ok = 0: v (nvec*16 B)
// exceeds lds_max or a row exceeds 16 waves x 9 vectors.
struct Sv64Shape {
    int ok, qv, wpr, blocks;
};
Sv64Shape nq_sv64_shape(long long nvec, long long S, size_t lds_max);
// raw sums: out[p] = sum_s conj(O_sp) w_s (p < P), out[P] = sum_s w_s, w_s = O_s^T v - ov
// (v, ov, out on the device, D = double or complex double); part: blocks * nvec * W D,
// wsum: blocks D, scratch: (blocks/32 + 1) * (P + 1) D
hipError_t nq_launch_sv_rows64(bool cplx, const Sv64Shape &sh, const void *O, const void *v, const void *ov,
                               long long S, long long nvec, long long P, void *part, void *wsum, void *scratch,
                               void *out, hipStream_t st);
// the static scratch of nqes_apply_S_rows_noalloc (noalloc.hip): resolves its device
// address on `device` (current) once; ptr / lds_max may be NULL
int nq_noalloc_setup(int device, void **ptr, int *lds_max);
const char *nq_noalloc_last_error();
// out = o^T v (one D), a fixed summation order
hipError_t nq_launch_dot64(bool cplx, const void *o, const void *v, long long P, void *out, hipStream_t st);
// out[p] = (raw[p] - conj(o_p) raw[P]) / S + skk_p v_p
hipError_t nq_launch_sv_finish64(bool cplx, const void *raw, const void *o, const void *skk, const void *v,
                                 long long P, double S, void *out, hipStream_t st);

// ---------------------------------------------------------------- mixed precision
// Arguments of the mixed sampler (G lanes per chain).  T = float or complex float
// (real / complex weights), D = double or complex double.  O rows have stride Pld
// (P rounded up to 16 bytes); H_loc and E_loc are complex double in both modes.
struct MixedSampleArgs {
    int N, M, NW, L;
    long long P, Pld;
    int C, m, B;
    double h;
    const double *J;           // N x N
    int model;                 // 0 ising-like, 1 / 2 hs-like with delta = 1 / -1
    double delta;
    const void *chT, *shT;     // K*N*M T: cosh(2w), sinh(2w) [kk][i][j]
    const void *wT, *bT;       // K*N*M T: w [kk][i][j]; K*M T: b [kk][j]
    const void *eaP, *eaN;     // K*N D: exp(2a), exp(-2a) [kk][i]
    const void *aD;            // K*N D: a [kk][i]
    const void *theta0;        // C*K*K*M T: angles at chain start [c][k][kk][j]
    const uint32_t *spins0;    // C*K*NW
    const uint64_t *rng0;      // 2*C pcg64 states after the chain-start draws
    void *O;                   // S*Pld T
    void *Hs;                  // S*K*K complex double
    void *Es;                  // S complex double
    uint32_t *spins;           // S*K*NW
    unsigned long long *accepts; // C
    unsigned int *flags;       // bit 0: an angle exceeded the FP32 range
    int jlds;                  // k_measure_mx with its tables in L2: J in LDS
    void *U;                   // (S*K) x Ucols T: S*v factors (NULL: not stored)
    int Ucols;
    float *Umax;               // S*K: max |U| of each factor row (real weights; NULL: not written)
};

// Factored S*v (sv_factored.hip): raw[p] = sum_s conj(O_sp) w_s, raw[P] = sum_s w_s,
// w_s = O_s^T v - ov, from the spins and the factor rows U (row r = s*K + kk:
// [k*M + j] = c_{s,k,kk} tanh(theta)_{kk,k,j}, [K*M + k] = c_{s,k,kk}, zero padded).
struct SvfArgs {
    int N, M, K, NW, Ucols;
    int n4;                    // K dimension of the forward MFMA: N rounded up to 4
    long long S, P, L;
    const uint32_t *spins;     // S*K*NW
    const void *U;             // (S*K) x Ucols T
    const void *v;             // P D
    const void *ov;            // 1 D
    void *wrow;                // S*K D
    void *w;                   // S D
    void *part;                // bwd_blocks x (P + 1) D
    const void *E;             // S complex double: local energies (moments)
    int wsel;                  // backward-pass weights: 0 w, 1 ones, 2 E
    int ncg;                   // forward column groups (wrow holds ncg x S*K partial sums)
    const void *Uq;            // int8 form: quantised factors (16 x 16 tiles of digit planes)
    const void *Uf;            // int8 form, GEMM forward pass: the same digits row-major per tile
    const double *escale;      // int8 form: per-row factor scale
    const float *Umax;         // int8 form: max |U| per factor row (from the measurement kernel)
    void *vq;                  // int8 form: v operands, column scales and biases
    void *q8w;                 // int8 form: backward block scales and weight digits
    void *q8m;                 // int8 form: spin masks of the backward A operand
    int qcols;                 // int8 form: digit columns per row (Ucols, or K M when the K c columns
                               // are kept apart as integers in xc)
    float *xc;                 // int8 form with qcols < Ucols: the c columns' quantised integers [rows][K]
};
// ok = 0: the factored form does not apply (N > 64).  The forward pass splits the
// Ucols/16 column tiles into fwd_cgs groups of fwd_cgt (B of a group fits in LDS); the
// backward pass into bwd_cgs groups of up to bwd_waves tiles, one wavefront per tile.
struct SvfShape {
    int ucols, ok, bwd_waves, bwd_blocks, bwd_cgs, fwd_blocks, fwd_cgt, fwd_cgs;
};
SvfShape nq_svf_shape(bool cplx, int N, int M, int K);
hipError_t nq_launch_svf_combine(bool cplx, const SvfArgs &a, hipStream_t st);
// |O|^2 sums and the packed moments, given the w = 1 and w = E sums at tmp, tmp + (P+1) D
hipError_t nq_launch_svf_sq_pack(bool cplx, const SvfArgs &a, const SvfShape &sh, void *part, void *scratch,
                                 void *tmp, void *so, void *ss, void *sf, hipStream_t st);
// int8 form of the factored S*v (real weights, N <= 64): factors quantised once per
// sample set (nq_q8_pack), S*v on int8 MFMA (nq_q8_apply)
struct Q8Shape {
    int ok, ucols, nsp, fwd_cgt, fwd_cgs, fwd_blocks, fwd_threads, bwd_waves, bwd_cgs, bwd_blocks;
    int gemm, ncg64, nst;  // GEMM forward pass (N <= 32): 64-column groups, spin tiles
    size_t fwd_lds, vq_bytes, w_bytes, m_bytes, uf_bytes, qg_bytes;
    long long rpad;  // rows of Uq / escale (multiple of 64)
    int gblk_ints;
    // one-pass form (sv_q8_fused.hip; real weights, N <= 32, K in {1, 2, 4}): the digits in
    // layout Ug (held in Uq), the quantised c columns in q8m, fz_blocks blocks of 64-row
    // chunks, fz_ncg 64-column groups of the K M tanh columns, fz_nst spin tiles
    int fused, fz_ncg, fz_nst, fz_blocks;
    size_t ug_bytes;  // layout Ug (3 digit planes x 64 fz_ncg columns per row)
    // two-pass GEMM form with K in {1, 2, 4}: the K c columns are kept apart (quantised
    // integers, FP64 arithmetic) when they would start a column tile of their own:
    // qcols = K M < ucols digit columns, xc_bytes per digit set (else qcols = ucols, 0)
    int qcols;
    size_t xc_bytes;
};
Q8Shape nq_q8_shape(int N, int M, int K, long long R, bool cplx = false);
// the one-pass form, per K (sv_q8_fused.hip compiled once per K)
hipError_t nq_q8f_pack_k1(const SvfArgs &a, const Q8Shape &sh, hipStream_t st);
hipError_t nq_q8f_pack_k2(const SvfArgs &a, const Q8Shape &sh, hipStream_t st);
hipError_t nq_q8f_pack_k4(const SvfArgs &a, const Q8Shape &sh, hipStream_t st);
hipError_t nq_q8f_apply_k1(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out, hipStream_t st);
hipError_t nq_q8f_apply_k2(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out, hipStream_t st);
hipError_t nq_q8f_apply_k4(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out, hipStream_t st);
hipError_t nq_q8f_moment_k1(const SvfArgs &a, const Q8Shape &sh, int wsel, void *part, void *scratch, void *out,
                            hipStream_t st);
hipError_t nq_q8f_moment_k2(const SvfArgs &a, const Q8Shape &sh, int wsel, void *part, void *scratch, void *out,
                            hipStream_t st);
hipError_t nq_q8f_moment_k4(const SvfArgs &a, const Q8Shape &sh, int wsel, void *part, void *scratch, void *out,
                            hipStream_t st);
hipError_t nq_q8_pack(const SvfArgs &a, const Q8Shape &sh, hipStream_t st);
hipError_t nq_q8_apply(const SvfArgs &a, const Q8Shape &sh, void *scratch, void *out, hipStream_t st);
// its forward pass as one GEMM over the columns (sh.gemm; sv_q8_gemm.hip): w_s directly for
// K = 1, 2, 4 unless rows, else the row sums wrow[r] (a.ncg = 1)
// nsets > 1 (K in {1, 2, 4}): that many digit sets in one launch, set z at a.Uf + z fss
// bytes, a.vq + z vss bytes (operands from nq_q8_prep_vg_sets), a.w + z wss doubles,
// a.xc + z xss floats
hipError_t nq_q8_fwd_gemm(const SvfArgs &a, const Q8Shape &sh, hipStream_t st, bool rows = false, bool prep = true,
                          int nsets = 1, long long fss = 0, long long vss = 0, long long wss = 0, long long xss = 0);
hipError_t nq_q8_prep_vg_sets(const SvfArgs &a, const Q8Shape &sh, int nsets, long long vss, size_t qss_bytes,
                              hipStream_t st);
// Complex weights on the int8 form (three real passes per product, Gauss's form): the
// complex factor rows a.U (complex float) are quantised with one scale per row into the
// digits of X_R = Re, X_I = Im and X_A = X_R + X_I (exact integers), three slices of
// a.Uq / a.Uf (slice t at t * sh.rpad * Ucols * 3 / t * sh.uf_bytes bytes), masks once.
hipError_t nq_q8c_pack(const SvfArgs &a, const Q8Shape &sh, hipStream_t st);
// per call: vin = v_A | v_R | v_I (P doubles each: Re v, Im v - Re v, Re v + Im v) | o^T v
// (re, im) | 0, 0; w3: 3 S doubles; out: the three backward runs' raw sums, (P + 1)
// doubles each (A, R, I); Re raw = A - R - I, Im raw = R - I (p < P), raw[P] = (I[P], R[P])
hipError_t nq_q8c_apply(const SvfArgs &a, const Q8Shape &sh, const double *vin, double *w3, void *scratch, double *out,
                        hipStream_t st);
void nq_q8c_sizes(const Q8Shape &sh, long long P, size_t *vq, size_t *q8w, size_t *part, size_t *scratch);
// moments from the quantised factors (sums with w = 1 and w = E on the int8 backward
// pass, |O|^2 on FP64 MFMA); after nq_q8_pack
hipError_t nq_q8_moments(const SvfArgs &a, const Q8Shape &sh, const SvfShape &svf, void *part, void *scratch,
                         void *tmp, void *so, void *ss, void *sf, hipStream_t st);
hipError_t nq_launch_svf(bool cplx, const SvfArgs &a, const SvfShape &sh, void *scratch, void *out, hipStream_t st);
// moments from the factors: so = sum_s O_s, ss = sum_s |O_s|^2, sf = sum_s E_s conj(O_s)
// (P complex double each); part: >= max(1, K(K-1)/2) * bwd_blocks * (P + 1) D,
// scratch: >= that / 32 + 2 (P + 1) D, tmp: 3 (P + 1) D
hipError_t nq_launch_svf_moments(bool cplx, const SvfArgs &a, const SvfShape &sh, void *part, void *scratch, void *tmp,
                                 void *so, void *ss, void *sf, hipStream_t st);
// out[col] = sum over chunks of part[chunk*stride + col] in chunk order (D = double or
// complex double); scratch >= (nchunk/32 + 1) * ncol D
hipError_t nq_chunk_reduce(bool cplx, const void *part, long long ncol, long long stride, int nchunk,
                           void *scratch, void *out, hipStream_t st);
// nsets sets of FP64 partial rows (set z at part + z pss, scratch: nsets ceil(nchunk / 32)
// ncol values) reduced into out + z oss, each in the order nq_chunk_reduce uses
hipError_t nq_chunk_reduce_sets(const double *part, long long ncol, long long stride, int nchunk, double *scratch,
                                double *out, int nsets, long long pss, long long oss, hipStream_t st);

hipError_t nq_launch_tables_mx(bool cplx, int N, int M, int K, int L, const void *W, void *chT, void *shT, void *wT,
                               void *bT, void *eaP, void *eaN, void *aD, hipStream_t st);
hipError_t nq_launch_chain_init(int N, int K, int NW, int C, uint64_t seed, uint32_t *spins0, uint64_t *rng0,
                                hipStream_t st);
hipError_t nq_launch_theta_gemm(bool cplx, int N, int M, int K, int NW, int C, const uint32_t *spins0,
                                const void *wT, const void *bT, void *theta0, hipStream_t st);
// Metropolis sweeps of the mixed path (sampler_mx.hip): G lanes per chain, JQ hidden
// units per lane (tables padded to G*JQ), wpb wavefronts per block, tables in LDS if tab.
struct SweepShape {
    int ok, G, JQ, wpb, tab;
    size_t tab_bytes;  // padded tables (device buffer passed to nq_launch_sweep_mx)
};
SweepShape nq_sweep_shape(bool cplx, int N, int M, int K, size_t lds_max);
template <int K>
hipError_t nq_sweep_k(bool cplx, const MixedSampleArgs &a, const void *tabp, const SweepShape &sh, hipStream_t st);
// Local energies and O / factor rows (sampler_mx.hip) from the sampler's padded tables
struct MeasureShape {
    int ok, G, JQ, wpb, tab;
    int jlds;  // tables in L2: J staged in LDS (it fits next to the per-sample states)
};
MeasureShape nq_measure_shape(bool cplx, int N, int K, const SweepShape &sw, size_t lds_max);
template <int K>
hipError_t nq_measure_k(bool cplx, const MixedSampleArgs &a, const void *tabp, const MeasureShape &sh, hipStream_t st);
hipError_t nq_launch_measure_mx(bool cplx, int K, const MixedSampleArgs &a, const MeasureShape &sh, const void *tabp,
                                hipStream_t st);
// tabp: the padded tables (sweeps); tabm: the measurement's copy (same size, cosh/sinh
// pairs spin-major); both written here
hipError_t nq_launch_sweep_mx(bool cplx, int K, const MixedSampleArgs &a, const SweepShape &sh, void *tabp,
                              void *tabm, hipStream_t st);

hipError_t nq_launch_moments_mx(bool cplx, const void *O, const void *E, long long S, long long P, long long Pld,
                                int nchunk, void *po, void *ps, void *pf, void *scratch, void *out_o, void *out_s,
                                void *out_f, hipStream_t st);
hipError_t nq_launch_sv_mx(bool cplx, const void *O, const void *v, const void *ov, long long S, long long P,
                           long long Pld, int nblk, void *part, void *wpart, void *wbuf, void *scratch, void *out,
                           hipStream_t st);
// launch shape of the one-pass S*v (qv = 0: two-pass form)
struct SvShape {
    int qv, wpr, np, blocks;
    size_t lds;
};
SvShape nq_sv_shape(bool cplx, long long nvec);
