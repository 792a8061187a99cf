// Dual numbers (value, tangent) for the tangent-linear MPDATA step: the
// per-stage kernels instantiated with Dual<double> compute the step and its
// directional derivative together.  Every branch (the signs in donor, pp,
// pn, the max/min selections of the limiter, |x|) is decided on the value,
// so the tangent is the derivative of the piecewise-smooth step on the
// branch the value takes; the adjoint kernels transpose exactly this
// linearisation.
#pragma once
#include "mpdata_common.hpp"

namespace mpdata {

// the scalar overloads stay visible beside the Dual ones below
using ::copysign;
using ::fabs;

template <typename R>
struct Dual {
    R v, d;
    Dual() = default;
    __host__ __device__ constexpr Dual(R a) : v(a), d(R(0)) {}
    __host__ __device__ constexpr Dual(R a, R b) : v(a), d(b) {}
};

template <typename T> struct is_dual { static constexpr bool value = false; };
template <typename R> struct is_dual<Dual<R>> { static constexpr bool value = true; };

template <> struct Eps<Dual<double>> { static constexpr Dual<double> v = Dual<double>(1.e-10); };

#define MPDATA_DUAL_FN template <typename R> __host__ __device__ __forceinline__
MPDATA_DUAL_FN Dual<R> operator+(Dual<R> a, Dual<R> b) { return {a.v + b.v, a.d + b.d}; }
MPDATA_DUAL_FN Dual<R> operator-(Dual<R> a, Dual<R> b) { return {a.v - b.v, a.d - b.d}; }
MPDATA_DUAL_FN Dual<R> operator-(Dual<R> a) { return {-a.v, -a.d}; }
MPDATA_DUAL_FN Dual<R> operator*(Dual<R> a, Dual<R> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
MPDATA_DUAL_FN Dual<R> operator/(Dual<R> a, Dual<R> b) {
    const R q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
MPDATA_DUAL_FN bool operator<(Dual<R> a, Dual<R> b) { return a.v < b.v; }
MPDATA_DUAL_FN bool operator>(Dual<R> a, Dual<R> b) { return a.v > b.v; }
MPDATA_DUAL_FN bool operator<=(Dual<R> a, Dual<R> b) { return a.v <= b.v; }
MPDATA_DUAL_FN bool operator>=(Dual<R> a, Dual<R> b) { return a.v >= b.v; }
// d|x| = sign(x) dx, with sign(0) = 0
MPDATA_DUAL_FN Dual<R> fabs(Dual<R> a) {
    return {a.v < R(0) ? -a.v : a.v, a.v > R(0) ? a.d : (a.v < R(0) ? -a.d : R(0))};
}
// SIGN(a, b) = |a| with the sign of b: the derivative follows a only
MPDATA_DUAL_FN Dual<R> copysign(Dual<R> a, Dual<R> b) {
    const R s = ::copysign(R(1), b.v) * ::copysign(R(1), a.v);
    return {::copysign(a.v, b.v), s * a.d};
}
#undef MPDATA_DUAL_FN

}  // namespace mpdata
