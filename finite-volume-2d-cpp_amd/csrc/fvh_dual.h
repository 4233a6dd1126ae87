// fvh_dual.h -- forward-mode dual numbers for local Jacobians.
//
// Every physics function in fvh_physics.h is a template on its number type,
// so instantiating it on Dual<T> gives the derivative of the branch the
// double-precision evaluation takes: comparisons (and so smin/smax, sabs and
// the limiter and boundary branches) act on the value part alone.  The
// adjoint uses this for the small per-boundary and per-cell Jacobians; the
// edge flux has a hand-written reverse (fvh_adjoint.h) that this type checks.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#ifndef FVH_HD
#define FVH_HD __host__ __device__ inline
#endif

namespace fvh {

// the overloads below must not hide the plain ones inside this namespace
using ::fabs;
using ::sqrt;

template <class T>
struct Dual {
    T v, d;
    FVH_HD Dual() : v(T(0)), d(T(0)) {}
    FVH_HD Dual(T a) : v(a), d(T(0)) {}   // constants (implicit, as literals appear in templates)
    FVH_HD Dual(T a, T b) : v(a), d(b) {}
};

template <class T> FVH_HD Dual<T> operator+(Dual<T> a, Dual<T> b) { return Dual<T>(a.v + b.v, a.d + b.d); }
template <class T> FVH_HD Dual<T> operator-(Dual<T> a, Dual<T> b) { return Dual<T>(a.v - b.v, a.d - b.d); }
template <class T> FVH_HD Dual<T> operator-(Dual<T> a) { return Dual<T>(-a.v, -a.d); }
template <class T> FVH_HD Dual<T> operator*(Dual<T> a, Dual<T> b) { return Dual<T>(a.v * b.v, a.d * b.v + a.v * b.d); }
template <class T> FVH_HD Dual<T> operator/(Dual<T> a, Dual<T> b) {
    const T r = a.v / b.v;
    return Dual<T>(r, (a.d - r * b.d) / b.v);
}
template <class T> FVH_HD bool operator<(Dual<T> a, Dual<T> b) { return a.v < b.v; }
template <class T> FVH_HD bool operator>(Dual<T> a, Dual<T> b) { return a.v > b.v; }
template <class T> FVH_HD bool operator<=(Dual<T> a, Dual<T> b) { return a.v <= b.v; }
template <class T> FVH_HD bool operator>=(Dual<T> a, Dual<T> b) { return a.v >= b.v; }
template <class T> FVH_HD bool operator==(Dual<T> a, Dual<T> b) { return a.v == b.v; }
template <class T> FVH_HD bool operator!=(Dual<T> a, Dual<T> b) { return a.v != b.v; }
template <class T> FVH_HD Dual<T> sqrt(Dual<T> a) {
    const T r = ::sqrt(a.v);
    return Dual<T>(r, a.d / (T(2) * r));
}
// |x| with the derivative sign taken from the value (+1 at 0)
template <class T> FVH_HD Dual<T> fabs(Dual<T> a) { return a.v < T(0) ? -a : a; }

}  // namespace fvh
