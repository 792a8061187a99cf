// The fused MPDATA step, exactly rounded: every operation of the reference in
// its order, rounded once (-ffp-contract=off, IEEE division), so the results
// are bit-identical to the per-stage kernels and the reference.  The kernel
// templates are in mpdata_fused_impl.hpp.
#include "mpdata_fused_impl.hpp"

namespace mpdata {

template <typename T>
bool Fused<T>::supports(const Geo& g, int variant, int iters, int nonos) {
    if (variant == 2) {
        // the standard kernel (mpdata_fused_std.hip): 2 iterations, limiter
        // on or off, any boundaries
        if (iters != 2 || (nonos != 0 && nonos != 1)) return false;
    } else {
        if (variant != 1 || (iters != 2 && iters != 3) || nonos != 1) return false;
        if (g.px || g.py || g.pz) return false;
    }
    // 32-bit element offsets inside the kernel, and 32-bit byte offsets
    // inside one plane of every array
    const long long biggest = len_of<FACEX>(g) > len_of<FACEY>(g) ? len_of<FACEX>(g) : len_of<FACEY>(g);
    const Off o = make_off(g);
    const long long plane = std::max(std::max((long long)o.sxy, (long long)o.fxy), (long long)o.yxy);
    return biggest < (1ll << 31) - 1 && len_of<FACEZ>(g) < (1ll << 31) - 1 &&
           plane * (long long)sizeof(T) < (1ll << 32);
}

template <typename T>
hipError_t FusedDens<T>::step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                              const T* u3, const T* h, const T* bcx, const T* bcy,
                              const DensIO<T>& d, hipStream_t s) {
    return dens_step<T, false>(g, x, xout, u1, u2, u3, h, bcx, bcy, d, s);
}

template <typename T>
const char* Fused<T>::scope() {
    return "the gauge variant with 2 or 3 iterations, the limiter on and open boundaries, "
           "and the standard variant with 2 iterations";
}

template <typename T>
hipError_t Fused<T>::step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                          const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                          hipStream_t s, int iters, const FusedScratch<T>* sc) {
    const Off o = make_off(g);
    hipError_t e = hipSuccess;
    int cfg = 0, cl = 64;
    choose_tiling<T>(g, cfg, cl);
    if (iters == 3) {
        if (!sc || !sc->x2 || !sc->ul[0] || !sc->ul[1] || !sc->ul[2]) return hipErrorInvalidValue;
        e = launch_three<T, false>(g, o, cfg, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, *sc, s);
        if (e != hipSuccess) return e;
        k_copy_untouched_halo<T><<<256, 256, 0, s>>>(g, x, xout);
        return hipGetLastError();
    }
    switch (cfg) {
        case 1: e = launch2<T, 32, 16, 4>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s); break;
        case 2: e = launch2<T, 32, 16, 2>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s); break;
        case 3: e = launch2<T, 64, 16, 4>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s); break;
        case 4: e = launch2<T, 16, 32, 2>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s); break;
        default: e = launch2<T, 32, 32, 4>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s); break;
    }
    if (e != hipSuccess) return e;
    k_copy_untouched_halo<T><<<256, 256, 0, s>>>(g, x, xout);
    return hipGetLastError();
}

// 32x16 tiles at 2 waves per SIMD (140 VGPRs, no spill), one workgroup
// per CU, and the chunk of least modelled wall time.  Alone the kernel is
// faster at 4 waves per SIMD (6 VGPRs spilled: 70.5 against 79.5 us per
// step at 128x128x64); but the adjoint's backward runs it beside the
// reverse kernels on another stream, and there the lighter occupancy
// leaves them room (forward + backward 16.3-16.5 against 16.7-16.9 ms).
template <typename T>
hipError_t Fused<T>::record(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                            const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                            const StdTape<T>& tp, hipStream_t s) {
    constexpr int RX = 32, RY = 16, MINB = 2;
    if (g.px || g.py || g.pz || g.ih != 1) return hipErrorInvalidValue;
    const Off o = make_off(g);
    const int env_clen = fused_chunk_env();
    int clen = env_clen > 0 ? env_clen : chunk_by_rounds<RX, RY>(g, k_fused_gauge_rec<T, RX, RY, MINB>);
    if (clen > g.lp) clen = g.lp;
    const int ntx = (g.np + RX - 5) / (RX - 4), nty = (g.mp + RY - 5) / (RY - 4);
    const int nchunk = (g.lp + clen - 1) / clen;
    int dev = 0, ncu = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_gauge_rec<T, RX, RY, MINB>,
                                                         RX * RY, 0);
    if (e != hipSuccess) return e;
    const int grid = persistent_grid((long long)ntx * nty * nchunk,
                                     (long long)(ntx * nty + NXCD - 1) / NXCD * nchunk, ncu, per_cu);
    // xout: the interior and the edge copies; the other halo cells are left
    // alone (the recorded step defines only the interior)
    k_fused_gauge_rec<T, RX, RY, MINB><<<grid, RX * RY, 0, s>>>(g, o, ntx, nty, nchunk, clen, x,
                                                                 u1, u2, u3, rhr, h, bcx, bcy,
                                                                 xout, tp);
    return hipGetLastError();
}

template struct Fused<double>;
template struct Fused<float>;
template struct FusedDens<double>;
template struct FusedDens<float>;

// ---- batched step ------------------------------------------------------------
// (tile configurations: MPDATA_BATCH_CONFIGS in mpdata_fused_impl.hpp)

template <typename T>
int FusedBatch<T>::nconfigs() {
    return 6;
}

template <typename T>
const char* FusedBatch<T>::config_name(int cfg) {
    switch (cfg) {
#define NAME(c, rx, ry, mb) \
    case c: return #rx "x" #ry " threads, " #mb " waves/SIMD";
        MPDATA_BATCH_CONFIGS(NAME)
#undef NAME
    }
    return "?";
}

template <typename T>
size_t FusedBatch<T>::table_bytes(int n) {
    return sched_offset<T>(n) + sizeof(BatchSched);
}

template <typename T>
int FusedBatch<T>::pack(const BatchMember<T>* m, int n, int cfg, int clen, void* table) {
    if (clen < 1) return -1;
    for (int q = 0; q < n; q++)
        if (!Fused<T>::supports(m[q].g, 1, 2, 1)) return -1;
    switch (cfg) {
#define PACK(c, rx, ry, mb) \
    case c: return pack_members<T, rx, ry>(m, n, clen, (Member<T>*)table);
        MPDATA_BATCH_CONFIGS(PACK)
#undef PACK
    }
    return -1;
}

template <typename T>
hipError_t FusedBatch<T>::step(const void* table, int n, int nitem, int cfg, long long max_halo,
                               hipStream_t s, int shared) {
    hipError_t e = hipErrorInvalidValue;
    switch (cfg) {
#define LAUNCH(c, rx, ry, mb) \
    case c: e = launch_batch<T, rx, ry, mb>(table, n, nitem, shared, s); break;
        MPDATA_BATCH_CONFIGS(LAUNCH)
#undef LAUNCH
    }
    if (e != hipSuccess) return e;
    const int bx = (int)std::min<long long>(64, (max_halo + 255) / 256);
    k_copy_untouched_halo_batch<T><<<dim3(std::max(bx, 1), n), 256, 0, s>>>((const Member<T>*)table);
    return hipGetLastError();
}

// Batch tiling for n members on one grid (configuration numbers of
// MPDATA_BATCH_CONFIGS): the single-domain rule with n times the work items.
template <typename T>
void FusedBatch<T>::choose_shared(const Geo& g, int n, int& cfg, int& clen) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
        ncu = 256;
    if (n * items_of<32, 32>(g, 64) >= 2LL * ncu) {
        cfg = 5;  // 32x32, 4 waves per SIMD
        clen = 64;
    } else {
        cfg = 1;  // 32x16, 4 waves per SIMD
        clen = 16;
        for (int c : {64, 32}) {
            if (n * items_of<32, 16>(g, c) >= 4LL * ncu) { clen = c; break; }
        }
    }
    if (clen > g.lp) clen = g.lp;
}

template struct FusedBatch<double>;
template struct FusedBatch<float>;

}  // namespace mpdata
