// The fused step of the standard antidiffusive variant, exactly rounded
// (-ffp-contract=off, IEEE division): bit-identical to the per-stage kernels.
// The kernel templates are in mpdata_fused_std_impl.hpp.
#include "mpdata_fused_std_impl.hpp"

namespace mpdata {

template <typename T>
hipError_t FusedStd<T>::step(const Geo& g, int nonos, const T* x, T* xout, const T* u1,
                             const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,
                             const T* bcy, hipStream_t s) {
    return std_step<T, false>(g, nonos, x, xout, u1, u2, u3, rhr, h, bcx, bcy, s);
}

template <typename T>
hipError_t FusedStd<T>::record(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                               const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                               const StdTape<T>& tp, hipStream_t s) {
    if (g.px || g.py || g.pz || g.ih != 1) return hipErrorInvalidValue;
    const Off o = make_off(g);
    int cfg = 1, cl = 64;
    choose_std(g, k_fused_std<T, 32, 16, 2, true, false, true>, cfg, cl);
    // xout: the interior and the edge copies; the other halo cells are left
    // alone (the recorded step defines only the interior)
    return launch_std<T, 32, 16, 2, true, false, true>(g, o, cl, x, xout, u1, u2, u3, rhr, h, bcx,
                                                       bcy, s, tp);
}

template struct FusedStd<double>;
template struct FusedStd<float>;

}  // namespace mpdata
