// sv_q8_common.hpp -- digit format shared by the int8 S*v kernels (sv_q8.hip,
// sv_q8_gemm.hip): balanced base-256 digits, least significant first.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int QU = 3;            // factor digits
constexpr int QV = 6;            // v digits
constexpr int QW = 6;            // weight digits
constexpr int NS = QU + QW - 1;  // digit weights 256^s of the backward products
// Digit pairs of weight 256^s with s < QSKIP are left out of both passes: (0, 0), (0, 1)
// and (1, 0) add at most 2^23 units to products of up to 2^69, 2^-46 of the largest, the
// size of the quantisation of v and w themselves (15 of the 18 MFMAs per digit set).
constexpr int QSKIP = 2;
constexpr double XMAX = 8355711.0;  // 127 (1 + 256 + 65536): |X| <= XMAX
constexpr int VBITS = 46;        // |Q| <= 2^46
constexpr int WBITS = 46;        // |Omega| <= 2^46
constexpr int MAX_BLOCK_CHUNKS = 256;  // 16384 rows: |int32 sums| <= 3 2^14 2^14 < 2^30

__device__ inline long long bal8l(long long x) { return (long long)(int8_t)(uint8_t)(x & 255); }

}  // namespace
