// kernels.hpp -- launch interface between the host library and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Arguments of the FP64 sampler (one wavefront per chain).  Complex arrays are
// interleaved doubles.  Sample s = chain * m + t.
struct Fp64SampleArgs {
    int N, M, NW, L;
    long long P;
    int C, m, B;
    uint64_t seed;
    double h;
    const double *J;           // N x N
    const void *W;             // P complex
    const void *chT, *shT;     // K*N*M complex: cosh(2w), sinh(2w), [kk][i][j]
    const void *eaP, *eaN;     // K*N complex: exp(2a), exp(-2a), [kk][i]
    void *O;                   // S*P complex (NULL: not stored)
    void *Hs;                  // S*K*K complex
    void *Es;                  // S complex: trace H_loc
    uint32_t *spins;           // S*K*NW
    unsigned long long *accepts; // C
};

hipError_t nq_launch_tables64(int N, int M, int K, int L, const void *W, void *chT, void *shT, void *eaP,
                              void *eaN, hipStream_t st);
hipError_t nq_launch_sample64(int K, const Fp64SampleArgs &a, hipStream_t st);
hipError_t nq_launch_moments64(const void *O, const void *E, long long S, long long P, int nchunk, void *po,
                               void *ps, void *pf, void *out_o, void *out_s, void *out_f, hipStream_t st);
hipError_t nq_launch_colsum64(const void *X, long long S, long long ncol, int nchunk, void *part, void *out,
                              hipStream_t st);
hipError_t nq_launch_sv64(const void *O, const void *v, long long S, long long P, int nchunk, void *w,
                          void *part, void *out, hipStream_t st);
hipError_t nq_launch_mathprobe64(const void *in, long long n, void *out, hipStream_t st);
