// dual.h - forward-mode dual numbers for the once-per-ray derivative pieces
// of the adjoint (camera frame, disc emission).  The per-step derivatives
// (Hamiltonian right-hand side, Dormand-Prince stages, dense output) are
// written out in reverse mode by hand in adjoint.hip; these cover the short
// closed-form chains at the ends of a ray, where N tangent directions cost
// less than a hand derivation is worth.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

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
    const double q = a.v / b.v;
    return scale_add(a, 1.0 / b.v, b, -q / b.v, q);
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
__host__ __device__ inline Dual<N> operator/(const Dual<N>& a, double b) { return chain(a, a.v / b, 1.0 / b); }
template <int N>
__host__ __device__ inline Dual<N> operator/(double b, const Dual<N>& a) {
    const double q = b / a.v;
    return chain(a, q, -q / a.v);
}

template <int N>
__host__ __device__ inline Dual<N> sqrt(const Dual<N>& a) {
    const double s = ::sqrt(a.v);
    return chain(a, s, 0.5 / s);
}
template <int N>
__host__ __device__ inline Dual<N> log(const Dual<N>& a) { return chain(a, ::log(a.v), 1.0 / a.v); }
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
    return chain(a, c, 1.0 / (3.0 * c * c));
}
template <int N>
__host__ __device__ inline Dual<N> pow(const Dual<N>& a, double p) {
    const double y = ::pow(a.v, p);
    return chain(a, y, p * y / a.v);
}

__host__ __device__ inline double value(double x) { return x; }
template <int N>
__host__ __device__ inline double value(const Dual<N>& x) { return x.v; }

}  // namespace bhx
