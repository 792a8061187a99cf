// Fused single-pass MPDATA step (mpdata_fused.hip).
#pragma once
#include "mpdata_common.hpp"

namespace mpdata {

// Device scratch of a three-iteration fused step: the second iterate and the
// limited velocities of the second pass (x-, y-, z-face arrays).
template <typename T>
struct FusedScratch {
    T* x2 = nullptr;
    T* ul[3] = {nullptr, nullptr, nullptr};
};

// Arguments of the pass modes of the fused kernel (PASS 1: the limited
// velocities written; PASS 2: the previous iterate read).
template <typename T>
struct PassIO {
    T *ul1 = nullptr, *ul2 = nullptr, *ul3 = nullptr;
    const T* xa = nullptr;
};

template <typename T>
struct StdTape;

template <typename T>
struct Fused {
    // True when a fused kernel implements this configuration: the gauge
    // variant with the limiter on, open walls and 2 or 3 iterations (this
    // struct's step), or the standard variant with 2 iterations, the limiter
    // on or off and any boundaries (FusedStd).
    static bool supports(const Geo& g, int variant, int iters, int nonos);
    // Human-readable scope of the fused kernel, for error messages.
    static const char* scope();
    // One full step out of place: xout <- MPDATA(x) with iters iterations
    // (3 needs sc).  xout ends up as the reference would leave x (new
    // interior, edge-copied x/y halo, other halo cells from x).  x and xout
    // must not overlap.
    static hipError_t step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                           const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                           hipStream_t s, int iters = 2, const FusedScratch<T>* sc = nullptr);
    // The two-pass step with open walls and ih = 1 that also records the
    // adjoint's tape (StdTape); xout receives the interior and the edge
    // copies, its other halo cells are left as they were.
    static hipError_t record(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                             const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                             const StdTape<T>& tp, hipStream_t s);
};

// The same fused gauge step built for speed rather than bit-identity: FMA
// contraction and Newton-refined reciprocal division (mpdata_fused_fast.hip).
// Each step differs from the exactly rounded one by a few units in the last
// place; after 50 steps of the 512x512x256 workload the normwise relative
// difference is of order 1e-14.  Same scope, layout and halo results as
// Fused.
template <typename T>
struct FusedFast {
    static hipError_t step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                           const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                           hipStream_t s, int iters = 2, const FusedScratch<T>* sc = nullptr);
};

// The fused step of the standard antidiffusive variant, 2 iterations, limiter
// on (nonos = 1) or off (0), open or periodic in each direction
// (mpdata_fused_std.hip).  Exactly rounded, bit-identical to the per-stage
// kernels; the same layout and halo results as Fused (periodic directions'
// halo cells keep their x values).
template <typename T>
struct FusedStd {
    static hipError_t step(const Geo& g, int nonos, const T* x, T* xout, const T* u1,
                           const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,
                           const T* bcy, hipStream_t s);
    // The limited step with open walls and ih = 1 that also records the
    // adjoint's tape (StdTape); xout receives the interior and the edge
    // copies, its other halo cells are left as they were.
    static hipError_t record(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                             const T* u3, const T* rhr, const T* h, const T* bcx, const T* bcy,
                             const StdTape<T>& tp, hipStream_t s);
};

// The same standard step built for speed rather than bit-identity
// (mpdata_fused_fast.hip): FMA contraction and Newton-refined reciprocal
// division, within 1e-12 of FusedStd; same scope, layout and halo results.
template <typename T>
struct FusedStdFast {
    static hipError_t step(const Geo& g, int nonos, const T* x, T* xout, const T* u1,
                           const T* u2, const T* u3, const T* rhr, const T* h, const T* bcx,
                           const T* bcy, hipStream_t s);
};

// Intermediates of a limited two-pass step (standard or gauge variant) that
// the adjoint step reads: x
// with the inflow values on its wall halo (xb, the upwind's input) and with
// edge copies there (x0, the limiter's bounds), the upwind result xant (psi)
// with its mirrored x/y halo ring, the pseudo-velocities V on the x-, y- and
// z-faces, and cp, cn.
template <typename T>
struct StdTape {
    T* xb = nullptr;
    T* x0 = nullptr;
    T* psi = nullptr;
    T* v[3] = {nullptr, nullptr, nullptr};
    T* cp = nullptr;
    T* cn = nullptr;
};

// Extra arrays of the density step (the reference's mpdata3d_rho_gauge_gpubc,
// advection/src_algorithms/mpdata3d-rho-gauge-gpubc.f90:23-622): x_incomp
// (read as the density at level n, written as x*h), pexe, the derived
// outputs, and the total mass fluxes m1..m3 (donor plus limited
// antidiffusive flux, on faces 1..n+1 of their direction).
template <typename T>
struct DensIO {
    T* xinc;
    const T* pexe;
    T *rhoadv, *rhr, *pexr1, *pexr2;
    T *m1, *m2, *m3;
};

// The density step after the velocity predictor in the fused kernel: the
// upwind divides by h and has no rhr factor, the update has none either,
// and the update plane also writes DensIO's outputs.  x_out / m1..m3 are
// out of place (x and u are read by neighbouring tiles' stencils); only
// the cells and faces the reference defines are written (interior and
// wall-line edge copies of x_out).
template <typename T>
struct FusedDens {
    static bool supports(const Geo& g) { return Fused<T>::supports(g, 1, 2, 1); }
    static hipError_t step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                           const T* u3, const T* h, const T* bcx, const T* bcy,
                           const DensIO<T>& d, hipStream_t s);
};

// The same density step built for speed (mpdata_fused_fast.hip: FMA
// contraction, Newton-refined reciprocal division), within 1e-12 of
// FusedDens; same scope and outputs.
template <typename T>
struct FusedDensFast {
    static hipError_t step(const Geo& g, const T* x, T* xout, const T* u1, const T* u2,
                           const T* u3, const T* h, const T* bcx, const T* bcy,
                           const DensIO<T>& d, hipStream_t s);
};

// One member of a batched step: its grid and fields (x in, xo out).
template <typename T>
struct BatchMember {
    Geo g;
    const T *x, *u1, *u2, *u3, *rhr, *h, *bcx, *bcy;
    T* xo;
};

// The fused step over many domains of different shapes in one persistent
// launch (plus one halo-copy launch).  The caller packs a descriptor table
// on the host (pack), copies it to device memory and launches with it.
template <typename T>
struct FusedBatch;

// The batched step of the fast build (mpdata_fused_fast.hip): the same
// tables, tilings and halo copy as FusedBatch, the fused item of FusedFast.
template <typename T>
struct FusedBatchFast {
    static hipError_t step(const void* table, int n, int nitem, int cfg, long long max_halo,
                           hipStream_t s, int shared = 0);
};

template <typename T>
struct FusedBatch {
    // number of tile configurations and a description of each
    static int nconfigs();
    static const char* config_name(int cfg);
    // bytes of the descriptor table for n members
    static size_t table_bytes(int n);
    // fills `table` (host memory, table_bytes(n)) for tile configuration cfg
    // and k-chunk length clen; returns the number of work items, or -1 if a
    // member is outside the fused kernel's scope or an argument is invalid
    static int pack(const BatchMember<T>* m, int n, int cfg, int clen, void* table);
    // one batched step with the device copy of a packed table; max_halo =
    // the largest number of halo cells of one member.  shared != 0: all
    // members have one grid, and the work items of one tile and k-chunk run
    // together for all members (fields sharing a velocity read it once)
    static hipError_t step(const void* table, int n, int nitem, int cfg, long long max_halo,
                           hipStream_t s, int shared = 0);
    // tile configuration and k-chunk for n members on grid g
    static void choose_shared(const Geo& g, int n, int& cfg, int& clen);
};

}  // namespace mpdata
