// The Courant maxima (common/diagutils.f90:588-672) without any allocation:
// the same two kernels as mpdata_compute_courants_*, with their per-workgroup
// partials in a file-scope device array instead of a context buffer, and the
// seven maxima written to the caller's device memory on the caller's stream.
// For callers that cannot allocate device memory: code under HIP graph
// capture, tight time loops, harnesses that forbid allocation.
//
// Every call on one GPU uses that GPU's single copy of the array: calls that
// overlap on one GPU (different streams, threads or processes' graphs that
// run concurrently) race.  Each call writes every partial it reads before
// reading it, so what a previous call left there never reaches a result.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mpdata_hip.h"
#include "mpdata_stages.hpp"
#include "mpdata_status.hpp"

using namespace mpdata;

namespace {

// Partials of at most kMaxBlocks workgroups of the per-cell grid (one
// workgroup per 64x4 columns of a plane: ceil(np/64)*ceil(mp/4)*lp), seven
// reals each, then the seven maxima: 3.67 MB per GPU.
constexpr long long kMaxBlocks = 65536;
static __device__ __align__(256) double scratch[(kMaxBlocks + 1) * 7];

// Device address of `scratch` per device.  In host code `scratch` names a
// host-side stand-in; its device address is looked up once per device, with
// that device current, and only that address is handed to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

int scratch_addr(void** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(MPDATA_EUNSUPPORTED, "courants_noalloc: device %d (0..%d supported)", dev,
                    kMaxDevices - 1);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(MPDATA_EHIP, "hipGetSymbolAddress(scratch) gave a null address");
        g_addr[dev] = p;
    }
    *out = g_addr[dev];
    return MPDATA_OK;
}

template <typename T>
__global__ void k_courants_out(const T* __restrict__ m, double* __restrict__ out) {
    if (threadIdx.x < 7) out[threadIdx.x] = (double)m[threadIdx.x];
}

template <typename T>
int courants_noalloc(const T* ox, const T* oy, const T* oz, const T* h, int level, T dxi, T dyi,
                     T dzi, T dt, int np, int mp, int lp, int ih, double* out, hipStream_t s) {
    if (!ox || !oy || !oz || !out || (level == 2 && !h)) return fail(MPDATA_EINVAL, "null pointer");
    if (level != 1 && level != 2)
        return fail(MPDATA_EINVAL, "ntimelevel %d (reference accepts 1 or 2)", level);
    if (np < 1 || mp < 1 || lp < 1 || ih < 1)
        return fail(MPDATA_EINVAL, "np,mp,lp,ih must be >= 1 (got %d,%d,%d,%d)", np, mp, lp, ih);
    const Geo g{np, mp, lp, ih, 0, 0, 0};
    const long long nb = Stages<T>::reduce_blocks(g);
    if (nb > kMaxBlocks)
        return fail(MPDATA_EINVAL,
                    "courants_noalloc: %dx%dx%d needs %lld reduction workgroups "
                    "(ceil(np/64)*ceil(mp/4)*lp); the static scratch holds %lld "
                    "(e.g. 256x256x256)",
                    np, mp, lp, nb, kMaxBlocks);
    void* p = nullptr;
    int rc = scratch_addr(&p);
    if (rc) return rc;
    T* part = (T*)p;
    HIPCHK(Stages<T>::courants(g, ox, oy, oz, h, level, dt * dxi, dt * dyi, dt * dzi, part, s));
    k_courants_out<T><<<1, 64, 0, s>>>(part + 7 * nb, out);
    HIPCHK(hipGetLastError());
    return MPDATA_OK;
}

}  // namespace

extern "C" {

int mpdata_compute_courants_noalloc_f64(const double* ox, const double* oy, const double* oz,
                                        const double* h, int ntimelevel, double dxi, double dyi,
                                        double dzi, double dt, int np, int mp, int lp, int ih,
                                        double* out, void* stream) {
    return courants_noalloc<double>(ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt, np, mp, lp, ih,
                                    out, (hipStream_t)stream);
}

int mpdata_compute_courants_noalloc_f32(const float* ox, const float* oy, const float* oz,
                                        const float* h, int ntimelevel, float dxi, float dyi,
                                        float dzi, float dt, int np, int mp, int lp, int ih,
                                        double* out, void* stream) {
    return courants_noalloc<float>(ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt, np, mp, lp, ih,
                                   out, (hipStream_t)stream);
}

}  // extern "C"
