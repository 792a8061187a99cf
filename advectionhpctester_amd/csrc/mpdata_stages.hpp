// Host-side launchers of the per-stage kernels (mpdata_stages.hip) and of
// the fused step (mpdata_fused.hip).
#pragma once
#include "mpdata_common.hpp"

namespace mpdata {

template <typename T>
struct Stages {
    // halo modes for halo_xy
    enum { MIRROR = 0, COPY = 1, ZERO = 2 };
    static hipError_t putbcinxh(const Geo& g, T* x, const T* bcx, const T* bcy, hipStream_t s);
    static hipError_t halo_xy(const Geo& g, T* a, int mode, int full, hipStream_t s);
    static hipError_t face_bc(const Geo& g, T* a1, T* a2, T* a3, hipStream_t s);
    static hipError_t upwind(const Geo& g, const T* u1, const T* u2, const T* u3, const T* x,
                             T* xant, const T* rhr, const T* h, hipStream_t s);
    static hipError_t pseudo_vel(const Geo& g, int variant, const T* u1, const T* u2,
                                 const T* u3, const T* xa, const T* h, T* v1, T* v2, T* v3,
                                 T* f1ns, T* f2ns, T* f3ns, hipStream_t s);
    static hipError_t limiter(const Geo& g, const T* x0, const T* xa, const T* h, const T* w1,
                              const T* w2, const T* w3, T* cp, T* cn, hipStream_t s);
    static hipError_t gauge_update(const Geo& g, int nonos, const T* v1, const T* v2,
                                   const T* v3, const T* cp, const T* cn, const T* psi,
                                   const T* rhr, const T* h, T* out, T* ul1, T* ul2, T* ul3,
                                   hipStream_t s);
    static hipError_t std_flux(const Geo& g, int nonos, const T* v1, const T* v2, const T* v3,
                               const T* cp, const T* cn, const T* psi, const T* rhr, T* f1,
                               T* f2, T* f3, T* ul1, T* ul2, T* ul3, hipStream_t s);
    static hipError_t std_update(const Geo& g, const T* f1, const T* f2, const T* f3,
                                 const T* psi, const T* rhr, const T* h, T* out, hipStream_t s);
    static hipError_t copy_interior(const Geo& g, const T* src, T* dst, hipStream_t s);
    static hipError_t velprd_extrap(const Geo& g, T* ox, T* oy, T* oz, const T* h, T gc1, T gc2,
                                    T gc3, T dt_ratio, hipStream_t s);
    static hipError_t a_to_c(const Geo& g, T* ox, T* oy, T* oz, T* u1, T* u2, T* u3,
                             hipStream_t s);
    // velprd_extrap + a_to_c in one pass without storing the A-grid level
    // n+1/2, level n-1 read as oxp*h (open walls; the caller rotates the time
    // levels by swapping the level-n and level-(n-1) arrays)
    static hipError_t predict_ac(const Geo& g, const T* ox, const T* oy, const T* oz,
                                 const T* oxp, const T* oyp, const T* ozp, const T* h, T gc1,
                                 T gc2, T gc3, T dt_ratio, T* u1, T* u2, T* u3, hipStream_t s);
    static hipError_t velprd_rotate(const Geo& g, T* u, T* v, T* w, T* ox, T* oy, T* oz,
                                    const T* h, T r, hipStream_t s);
    static long long reduce_blocks(const Geo& g);
    // Courant maxima: per-block partials in part[0 .. 7*reduce_blocks), then
    // the seven maxima in part[7*reduce_blocks .. +7)
    static hipError_t courants(const Geo& g, const T* ox, const T* oy, const T* oz, const T* h,
                               int level, T gc1, T gc2, T gc3, T* part, hipStream_t s);
    static hipError_t norm_sums(const Geo& g, const T* x, const T* x0, double* part,
                                hipStream_t s);
};

}  // namespace mpdata
