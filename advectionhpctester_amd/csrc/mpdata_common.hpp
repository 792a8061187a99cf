// Shared device-side definitions of the MPDATA kernels for gfx950.
//
// Field layout (the reference's, unchanged): Fortran column-major arrays with
// lower bound 1-ih in each direction, i fastest.
//   scalar  x,xant,h,rhr,cp,cn : (np+2ih)   x (mp+2ih)   x (lp+2ih)
//   x-faces u1,v1,f1,f1ns      : (np+1+2ih) x (mp+2ih)   x (lp+2ih)
//   y-faces u2,v2,f2,f2ns      : (np+2ih)   x (mp+1+2ih) x (lp+2ih)
//   z-faces u3,v3,f3,f3ns      : (np+2ih)   x (mp+2ih)   x (lp+1+2ih)
// (advection/src_algorithms/upwind3d-gpubc.f90:59-78,
//  common/scratch_datafields_modular.f90:282-306)
//
// Every expression keeps the reference's operand order.  The library is
// compiled with -ffp-contract=off and IEEE division so that each operation
// rounds once, as in the reference CPU build.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpdata {

// Grid geometry and boundary kinds.  ibc = 0 open (reference), 1 periodic
// (clean period-n wrap in all three directions; build-defined).
struct Geo {
    int np, mp, lp, ih;
    int px, py, pz;  // periodic flags
};

enum Kind { SCAL = 0, FACEX = 1, FACEY = 2, FACEZ = 3 };

__host__ __device__ inline int wrap(int i, int n) { return ((i - 1) % n + n) % n + 1; }

template <int K>
__host__ __device__ inline long long nx_of(const Geo& g) {
    return g.np + 2 * g.ih + (K == FACEX);
}
template <int K>
__host__ __device__ inline long long ny_of(const Geo& g) {
    return g.mp + 2 * g.ih + (K == FACEY);
}
template <int K>
__host__ __device__ inline long long nz_of(const Geo& g) {
    return g.lp + 2 * g.ih + (K == FACEZ);
}
template <int K>
__host__ __device__ inline long long len_of(const Geo& g) {
    return nx_of<K>(g) * ny_of<K>(g) * nz_of<K>(g);
}

// Offset of Fortran element (i,j,k) without wrapping.
template <int K>
__host__ __device__ inline long long off_raw(const Geo& g, int i, int j, int k) {
    return (long long)(i + g.ih - 1) +
           nx_of<K>(g) * ((long long)(j + g.ih - 1) + ny_of<K>(g) * (long long)(k + g.ih - 1));
}

// Offset with indices of periodic directions wrapped to 1..n (face n+1 is
// face 1 of the same period).
template <int K>
__host__ __device__ inline long long off(const Geo& g, int i, int j, int k) {
    if (g.px) i = wrap(i, g.np);
    if (g.py) j = wrap(j, g.mp);
    if (g.pz) k = wrap(k, g.lp);
    return off_raw<K>(g, i, j, k);
}

// ---- point operators: advection/src_algorithms/mpdataoperators.inc:1-96 ----
// XCDs of an MI355X; the workgroups of a launch are dealt to them
// round-robin (workgroup b to XCD b mod NXCD), each XCD with its own L2.
constexpr int NXCD = 8;

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double v = 1.e-10; };  // epsilons.f90:4
template <> struct Eps<float> { static constexpr float v = 1.e-10f; };

template <typename T> __device__ __forceinline__ T pp(T y) { return y > T(0) ? y : T(0); }
template <typename T> __device__ __forceinline__ T mn0(T y) { return y < T(0) ? y : T(0); }
template <typename T> __device__ __forceinline__ T pn(T y) { return -mn0(y); }
template <typename T> __device__ __forceinline__ T donor(T y1, T y2, T a) {
    return pp(a) * y1 - (-(mn0(a) * y2));
}
template <typename T> __device__ __forceinline__ T rat2(T z1, T z2) { return (z2 - z1) * T(0.5); }
template <typename T> __device__ __forceinline__ T rat4(T z0, T z1, T z2, T z3) {
    return ((z3 + z2) - (z1 + z0)) * T(0.25);
}
template <typename T> __device__ __forceinline__ T vdyf(T x1, T x2, T a, T rinv) {
    return (fabs(a) - a * a * rinv) * rat2(x1, x2);
}
template <typename T> __device__ __forceinline__ T vdyfa(T x1, T x2, T a, T rinv) {
    return (fabs(a) - a * a * rinv) * (fabs(x2) - fabs(x1)) / (fabs(x2) + fabs(x1) + Eps<T>::v);
}
template <typename T> __device__ __forceinline__ T vcorra(T a, T b, T y1, T y2, T rinv) {
    return -(T(0.125) * a * b * y1 * rinv / (y2));
}
template <typename T> __device__ __forceinline__ T min2(T a, T b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T max2(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T min3(T a, T b, T c) { return min2(min2(a, b), c); }

// Hardware max/min (v_max_f64 / v_min_f64: one instruction instead of a
// compare and two selects).  For finite operands they return the same value
// as the selects above except possibly the sign of a zero result.  The fused
// kernel uses them where a signed zero cannot reach the new field: every
// zero that max(0,v) or min(0,v) produces is multiplied into a flux, the flux
// sum is subtracted from a field value, and x - (+-0) = x in round-to-nearest;
// limiter bounds and ratios are max/min/sums of finite values none of which
// is -0 (a field value is -0 only if the input holds -0).
namespace hw {
template <typename T> __device__ __forceinline__ T max2(T a, T b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T min2(T a, T b) { return fmin(a, b); }
template <typename T> __device__ __forceinline__ T min3(T a, T b, T c) { return fmin(fmin(a, b), c); }
template <typename T> __device__ __forceinline__ T pp(T y) { return fmax(T(0), y); }
template <typename T> __device__ __forceinline__ T pn(T y) { return -fmin(T(0), y); }
template <typename T> __device__ __forceinline__ T donor(T y1, T y2, T a) {
    return fmax(T(0), a) * y1 - (-(fmin(T(0), a) * y2));
}
template <typename T>
__device__ __forceinline__ T gflux(T v, T cp_c, T cn_m, T cp_m, T cn_c) {
    return pp(v) * min3(T(1), cp_c, cn_m) - pn(v) * min3(T(1), cp_m, cn_c);
}
}  // namespace hw

// Gauge limited flux through one face (antidiff3d-gauge-gpubc.f90:405-422):
// v the pseudo-velocity, cp_c/cn_c at the downstream cell, cp_m/cn_m upstream.
template <typename T>
__device__ __forceinline__ T gflux(T v, T cp_c, T cn_m, T cp_m, T cn_c) {
    return pp(v) * min3(T(1), cp_c, cn_m) - pn(v) * min3(T(1), cp_m, cn_c);
}

// Sign-aware limited velocity of the standard variant
// (antidiff3d-standard-gpubc.f90:372-383); Fortran SIGN(1,y) is copysign.
template <typename T>
__device__ __forceinline__ T slimv(T v, T a1, T a2, T xm, T xc) {
    T sm = copysign(T(1), xm), smn = copysign(T(1), -xm);
    T sc = copysign(T(1), xc), scn = copysign(T(1), -xc);
    return pp(v) * (a1 * pp(sm) + a2 * pp(smn)) - pn(v) * (a2 * pp(sc) + a1 * pp(scn));
}

}  // namespace mpdata
