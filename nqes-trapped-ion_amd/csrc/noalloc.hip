// noalloc.hip -- the S*v entry point that allocates nothing (nqes_apply_S_rows_noalloc).
//
// The FP64 one-pass S*v (k_sv_rows64 and its ordered reductions, kernels_fp64.hip)
// needs device memory beyond its inputs and output: per-block partial rows, per-block
// sums of w, the reduction's group sums, the raw result and o^T v.  nqes_apply_S takes
// them from buffers its context allocated; this entry point takes them from one
// file-scope device array, `scratch`, so a caller that cannot call hipMalloc (code
// under HIP graph capture, a serving loop, a harness that forbids it) can still apply
// the operator to O rows it holds in device memory.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device
// address comes from hipGetSymbolAddress, looked up once per device with that device
// current (nq_noalloc_setup, called by nqes_create and by the entry point's first
// call on a device) and kept per device; kernels receive that address.  Every call on
// a device shares the device's array, so calls that overlap on one device race.  No
// region is read before the same call's kernels write it, so nothing is cleared.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nqes.h"
#include "kernels.hpp"

// the largest scratch the one-pass kernel's shapes need is ~20 MB (at most 2048 x 64 x 9
// partial-row vectors of 16 bytes, plus the reduction's group sums)
#define NQ_NOALLOC_SCRATCH_BYTES (24u << 20)

static __device__ __align__(256) unsigned char scratch[NQ_NOALLOC_SCRATCH_BYTES];

namespace {

struct DevScratch {
    void *ptr = nullptr;  // device address of `scratch` on this device
    int lds_max = 0;      // LDS bytes per block
};

std::mutex g_mu;
std::vector<DevScratch> g_dev;
thread_local std::string g_err;

int fail(int code, const char *fmt, long long a = 0, long long b = 0)
{
    char buf[256];
    snprintf(buf, sizeof(buf), fmt, a, b);
    g_err = buf;
    return code;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

} // namespace

// Resolves this device's `scratch` address (with `device` current).  NQES_OK or
// NQES_EDEVICE.
int nq_noalloc_setup(int device, void **ptr, int *lds_max)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0) return NQES_EDEVICE;
    if ((size_t)device >= g_dev.size()) g_dev.resize((size_t)device + 1);
    DevScratch &d = g_dev[(size_t)device];
    if (!d.ptr) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return NQES_EDEVICE;
        }
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 65536;
        d.ptr = p;
        d.lds_max = lds;
    }
    if (ptr) *ptr = d.ptr;
    if (lds_max) *lds_max = d.lds_max;
    return NQES_OK;
}

const char *nq_noalloc_last_error() { return g_err.c_str(); }

extern "C" int nqes_apply_S_rows_noalloc(const double *O, int64_t S, int64_t P, int64_t ldo, int cplx, const double *o,
                                         const double *skk, const double *v, double *out, void *stream)
{
    if (!O || !o || !skk || !v || !out) return fail(NQES_EINVAL, "nqes_apply_S_rows_noalloc: NULL pointer");
    if (S < 1 || P < 1) return fail(NQES_EINVAL, "nqes_apply_S_rows_noalloc: S=%lld P=%lld must be >= 1", S, P);
    if (cplx != 0 && cplx != 1) return fail(NQES_EINVAL, "nqes_apply_S_rows_noalloc: cplx=%lld is not 0 or 1", cplx);
    if (ldo < P || (!cplx && (ldo & 1)))
        return fail(NQES_EINVAL, "nqes_apply_S_rows_noalloc: row stride %lld must be >= P=%lld (and even for real O)",
                    ldo, P);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(NQES_EDEVICE, "nqes_apply_S_rows_noalloc: no current device");
    void *base = nullptr;
    int lds_max = 0;
    if (nq_noalloc_setup(dev, &base, &lds_max) != NQES_OK)
        return fail(NQES_EDEVICE, "nqes_apply_S_rows_noalloc: scratch address lookup failed on device %lld", dev);
    const long long nvec = cplx ? ldo : ldo / 2;  // 16-byte vectors per row
    const Sv64Shape sh = nq_sv64_shape(nvec, S, (size_t)lds_max);
    if (!sh.ok)
        return fail(NQES_EINVAL, "nqes_apply_S_rows_noalloc: rows of %lld 16-byte vectors exceed the one-pass kernel "
                                 "(v in %lld bytes of LDS, at most 16 x 64 x 9 vectors)",
                    nvec, lds_max);
    const size_t dsz = cplx ? 16 : 8;
    // layout: o^T v | raw (P + 1) | wsum (blocks) | partial rows | group sums
    const size_t off_raw = 256;
    const size_t off_wsum = off_raw + align256((size_t)(P + 1) * dsz);
    const size_t off_part = off_wsum + align256((size_t)sh.blocks * dsz);
    const size_t off_grp = off_part + align256((size_t)sh.blocks * nvec * 16);
    const size_t need = off_grp + align256((size_t)(sh.blocks / 32 + 2) * (P + 1) * dsz);
    if (need > NQ_NOALLOC_SCRATCH_BYTES)
        return fail(NQES_EINVAL, "nqes_apply_S_rows_noalloc: needs %lld scratch bytes, the static array holds %lld",
                    (long long)need, (long long)NQ_NOALLOC_SCRATCH_BYTES);
    char *b = static_cast<char *>(base);
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = nq_launch_dot64(cplx != 0, o, v, P, b, st);
    if (e == hipSuccess)
        e = nq_launch_sv_rows64(cplx != 0, sh, O, v, b, S, nvec, P, b + off_part, b + off_wsum, b + off_grp,
                                b + off_raw, st);
    if (e == hipSuccess) e = nq_launch_sv_finish64(cplx != 0, b + off_raw, o, skk, v, P, (double)S, out, st);
    if (e != hipSuccess) {
        g_err = std::string("nqes_apply_S_rows_noalloc: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? NQES_ENOMEM : NQES_EDEVICE;
    }
    return NQES_OK;
}
