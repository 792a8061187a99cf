// The fused steps of the fast build: the kernel templates of
// mpdata_fused_impl.hpp (gauge, density) and mpdata_fused_std_impl.hpp
// (standard)
// with FAST = true (Newton-refined reciprocal division), and this file
// compiled with -ffp-contract=fast so that products and sums contract into
// FMAs.  Not bit-identical to the reference's operation order; the exact
// builds are mpdata_fused.hip and mpdata_fused_std.hip.
#include "mpdata_fused_impl.hpp"
#include "mpdata_fused_std_impl.hpp"

namespace mpdata {

template <typename T>
hipError_t FusedFast<T>::step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                              const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                              hipStream_t s, int iters, const FusedScratch<T>* sc) {
    const Off o = make_off(g);
    int cfg = 0, cl = 64;
    choose_tiling<T>(g, cfg, cl);
    if (iters == 3) {
        if (!sc || !sc->x2 || !sc->ul[0] || !sc->ul[1] || !sc->ul[2]) return hipErrorInvalidValue;
        const hipError_t e =
            launch_three<T, true>(g, o, cfg, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, *sc, s);
        if (e != hipSuccess) return e;
        k_copy_untouched_halo<T><<<256, 256, 0, s>>>(g, x, xout);
        return hipGetLastError();
    }
    // the two tilings the chooser picks (32x32 lanes, 32x16 lanes at four
    // waves per SIMD); the tuning overrides of the exact build are not built
    const hipError_t e =
        cfg == 0 ? launch2<T, 32, 32, 4, true>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s)
                 : launch2<T, 32, 16, 4, true>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
    if (e != hipSuccess) return e;
    k_copy_untouched_halo<T><<<256, 256, 0, s>>>(g, x, xout);
    return hipGetLastError();
}

template <typename T>
hipError_t FusedBatchFast<T>::step(const void* table, int n, int nitem, int cfg,
                                   long long max_halo, hipStream_t s, int shared) {
    hipError_t e = hipErrorInvalidValue;
    switch (cfg) {
#define LAUNCH(c, rx, ry, mb) \
    case c: e = launch_batch<T, rx, ry, mb, true>(table, n, nitem, shared, s); break;
        MPDATA_BATCH_CONFIGS(LAUNCH)
#undef LAUNCH
    }
    if (e != hipSuccess) return e;
    const int bx = (int)std::min<long long>(64, (max_halo + 255) / 256);
    k_copy_untouched_halo_batch<T><<<dim3(std::max(bx, 1), n), 256, 0, s>>>((const Member<T>*)table);
    return hipGetLastError();
}

template <typename T>
hipError_t FusedStdFast<T>::step(const Geo& g, int nonos, const T* x, T* xout, const T* u1,
                                 const T* u2, const T* u3, const T* rhr, const T* h,
                                 const T* bcx, const T* bcy, hipStream_t s) {
    return std_step<T, true>(g, nonos, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
}

template <typename T>
hipError_t FusedDensFast<T>::step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                                  const T* u3, const T* h, const T* bcx, const T* bcy,
                                  const DensIO<T>& d, hipStream_t s) {
    return dens_step<T, true>(g, x, xout, u1, u2, u3, h, bcx, bcy, d, s);
}

template struct FusedFast<double>;
template struct FusedFast<float>;
template struct FusedDensFast<double>;
template struct FusedDensFast<float>;
template struct FusedStdFast<double>;
template struct FusedStdFast<float>;
template struct FusedBatchFast<double>;
template struct FusedBatchFast<float>;

}  // namespace mpdata
