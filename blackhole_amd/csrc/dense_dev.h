// dense_dev.h - the Dormand-Prince dense output and the event roots on it,
// shared by the tracer and the adjoint (host-callable for the adjoint's
// self-checks).
#pragma once

#include "kerr_dev.h"

namespace bhx {

__host__ __device__ __forceinline__ double dense_comp(double c1, double c2, double c3, double c4, double c5,
                                             double s) {
    const double s1 = 1.0 - s;
    return c1 + s * (c2 + s1 * (c3 + s * (c4 + s1 * c5)));
}

// cos(th) * c0 <= 0 for c0 = cos of a double (never 0): the sign of cos(th)
// is read off th near the equator, where cos(fl(pi/2)) = 6.1e-17 > 0 and the
// next double up has a negative cosine; elsewhere cos is evaluated.
__host__ __device__ __forceinline__ bool crossed(double th, double c0) {
    if (th > 0.5 && th < 2.6) {
        const bool cpos = th <= 1.5707963267948966;
        return (c0 > 0.0) ? !cpos : cpos;
    }
    double s, c;
    sincos_near(th, &s, &c);
    return c * c0 <= 0.0;
}

// Value and s-derivative of one component of the dense-output polynomial.
__host__ __device__ __forceinline__ void dense_vd(double c1, double c2, double c3, double c4, double c5,
                                         double s, double& v, double& dv) {
    const double s1 = 1.0 - s;
    const double h4 = c4 + s1 * c5;
    const double g = c3 + s * h4;
    const double dg = h4 - s * c5;
    const double m = c2 + s1 * g;
    const double dm = s1 * dg - g;
    v = c1 + s * m;
    dv = m + s * dm;
}

// Event roots on the dense output.  The reference bisects 60 times on the
// predicate (cos(theta(s)) c0 <= 0 for the disc, r(s) <= R for the star) and
// returns the midpoint of the last bracket (render.cpp:160-164, 184-190).
// Here the same predicate keeps a bracket while Newton's method on the
// smooth function behind it (theta(s) - fl(pi/2), r(s) - R) converges
// quadratically from the secant estimate: the same root to rounding in ~4
// iterations instead of 60, so one lane's event no longer stalls its wave.
// Bisection steps are taken whenever Newton leaves the bracket.
template <typename Pred>
__host__ __device__ __forceinline__ double newton_root(double c1, double c2, double c3, double c4, double c5,
                                              double target, double f0, double f1, Pred hit) {
    double lo = 0.0, hi = 1.0;
    double s = (f0 != f1) ? f0 / (f0 - f1) : 0.5;
    if (!(s > 0.0 && s < 1.0)) s = 0.5;
    for (int it = 0; it < 12; ++it) {
        double v, dv;
        dense_vd(c1, c2, c3, c4, c5, s, v, dv);
        if (hit(v)) hi = s; else lo = s;
        double sn = s - (v - target) / dv;
        if (sn == s) break;   // on the root to rounding (e.g. v == target exactly)
        if (!(sn >= lo && sn <= hi)) sn = 0.5 * (lo + hi);
        const bool done = fabs(sn - s) <= 4e-16;
        s = sn;
        if (done) break;
    }
    return s;
}

// theta(s) crosses the equatorial plane; c0 = cos(theta(0)).
__host__ __device__ __forceinline__ double root_theta(double c1, double c2, double c3, double c4, double c5,
                                             double c0, double th0, double th1) {
    if (th0 > 0.5 && th0 < 2.6 && th1 > 0.5 && th1 < 2.6) {
        const double T = 1.5707963267948966;
        return newton_root(c1, c2, c3, c4, c5, T, th0 - T, th1 - T,
                           [c0](double th) { return crossed(th, c0); });
    }
    double lo = 0.0, hi = 1.0;   // crossing through a pole's image: plain bisection
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (crossed(dense_comp(c1, c2, c3, c4, c5, mid), c0)) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// r(s) reaches the neutron-star surface R.
__host__ __device__ __forceinline__ double root_r(double c1, double c2, double c3, double c4, double c5,
                                         double R, double r0, double r1) {
    return newton_root(c1, c2, c3, c4, c5, R, r0 - R, r1 - R, [R](double r) { return r <= R; });
}

}  // namespace bhx
