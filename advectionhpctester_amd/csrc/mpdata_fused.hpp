// Fused single-pass MPDATA step (mpdata_fused.hip).
#pragma once
#include "mpdata_common.hpp"

namespace mpdata {

template <typename T>
struct Fused {
    // True when the fused kernel implements this configuration.
    static bool supports(const Geo& g, int variant, int iters, int nonos);
    // Human-readable scope of the fused kernel, for error messages.
    static const char* scope();
    // One full step out of place: xout <- MPDATA(x).  xout ends up as the
    // reference would leave x (new interior, edge-copied x/y halo, other
    // halo cells from x).  x and xout must not overlap.
    static hipError_t step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                           const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                           hipStream_t s);
};

}  // namespace mpdata
