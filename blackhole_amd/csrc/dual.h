// dual.h - forward-mode dual numbers for the closed-form derivative pieces of
// the adjoint (camera frame, disc emission, the conical jet's integrand at
// each quadrature node).  The per-step derivatives (Hamiltonian right-hand
// side, Dormand-Prince stages, dense output) are written out in reverse mode
// by hand in adjoint.hip; these cover short closed-form chains, where N
// tangent directions cost less than a hand derivation is worth.  A quotient
// takes one reciprocal (hardware estimate + Newton, <= 1 ulp) for its value
// and its derivative instead of the IEEE division sequence per term.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "kerr_dev.h"

namespace bhx {

// The overloads below must not hide the scalar functions for code in bhx.
using ::acos;
using ::cbrt;
using ::cos;
using ::exp;
using ::log;
using ::pow;
using ::sin;
using ::sqrt;

template <int N>
struct Dual {
    double v;
    double d[N];
    __host__ __device__ Dual() : v(0.0) {
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    __host__ __device__ Dual(double x) : v(x) {  // NOLINT: constants promote implicitly
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    __host__ __device__ static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

template <int N>
__host__ __device__ inline Dual<N> scale_add(const Dual<N>& a, double fa, const Dual<N>& b, double fb,
                                             double v) {
    Dual<N> r;
    r.v = v;
    for (int i = 0; i < N; ++i) r.d[i] = fa * a.d[i] + fb * b.d[i];
    return r;
}
template <int N>
__host__ __device__ inline Dual<N> chain(const Dual<N>& a, double v, double dv) {
    Dual<N> r;
    r.v = v;
    for (int i = 0; i < N; ++i) r.d[i] = dv * a.d[i];
    return r;
}

template <int N>
__host__ __device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    return scale_add(a, 1.0, b, 1.0, a.v + b.v);
}
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    return scale_add(a, 1.0, b, -1.0, a.v - b.v);
}
template <int N>
__host__ __device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    return scale_add(a, b.v, b, a.v, a.v * b.v);
}
template <int N>
__host__ __device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    const double ib = rcp(b.v);
    const double q = a.v * ib;
    return scale_add(a, ib, b, -q * ib, q);
}
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a) {
    return chain(a, -a.v, -1.0);
}
template <int N>
__host__ __device__ inline Dual<N> operator+(const Dual<N>& a, double b) { return chain(a, a.v + b, 1.0); }
template <int N>
__host__ __device__ inline Dual<N> operator+(double b, const Dual<N>& a) { return chain(a, a.v + b, 1.0); }
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a, double b) { return chain(a, a.v - b, 1.0); }
template <int N>
__host__ __device__ inline Dual<N> operator-(double b, const Dual<N>& a) { return chain(a, b - a.v, -1.0); }
template <int N>
__host__ __device__ inline Dual<N> operator*(const Dual<N>& a, double b) { return chain(a, a.v * b, b); }
template <int N>
__host__ __device__ inline Dual<N> operator*(double b, const Dual<N>& a) { return chain(a, a.v * b, b); }
template <int N>
__host__ __device__ inline Dual<N> operator/(const Dual<N>& a, double b) {
    const double ib = rcp(b);
    return chain(a, a.v * ib, ib);
}
template <int N>
__host__ __device__ inline Dual<N> operator/(double b, const Dual<N>& a) {
    const double ia = rcp(a.v);
    const double q = b * ia;
    return chain(a, q, -q * ia);
}

template <int N>
__host__ __device__ inline Dual<N> sqrt(const Dual<N>& a) {
    const double s = ::sqrt(a.v);
    return chain(a, s, 0.5 * rcp(s));
}
template <int N>
__host__ __device__ inline Dual<N> log(const Dual<N>& a) { return chain(a, ::log(a.v), rcp(a.v)); }
template <int N>
__host__ __device__ inline Dual<N> exp(const Dual<N>& a) {
    const double e = ::exp(a.v);
    return chain(a, e, e);
}
template <int N>
__host__ __device__ inline Dual<N> sin(const Dual<N>& a) { return chain(a, ::sin(a.v), ::cos(a.v)); }
template <int N>
__host__ __device__ inline Dual<N> cos(const Dual<N>& a) { return chain(a, ::cos(a.v), -::sin(a.v)); }
template <int N>
__host__ __device__ inline Dual<N> acos(const Dual<N>& a) {
    return chain(a, ::acos(a.v), -1.0 / ::sqrt(1.0 - a.v * a.v));
}
template <int N>
__host__ __device__ inline Dual<N> cbrt(const Dual<N>& a) {
    const double c = ::cbrt(a.v);
    return chain(a, c, rcp(3.0 * c * c));
}
template <int N>
__host__ __device__ inline Dual<N> pow(const Dual<N>& a, double p) {
    const double y = ::pow(a.v, p);
    return chain(a, y, p * y * rcp(a.v));
}

__host__ __device__ inline double value(double x) { return x; }
template <int N>
__host__ __device__ inline double value(const Dual<N>& x) { return x.v; }

}  // namespace bhx
