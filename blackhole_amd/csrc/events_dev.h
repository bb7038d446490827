// events_dev.h - optically thin jet emission and the Sun, on the dense output.
//
// Device forms of the reference's Jet (src/jet.h:156-248, 314), the
// sample_jet quadrature (src/render.cpp:47-50, 107-122), the segment cull
// (src/render.cpp:55-62, 151-152) and the Sun event (src/render.cpp:202-241).
// They run only on steps where the scene has a jet or a Sun, so they read
// their parameters through a pointer (scalar loads inside the branch) rather
// than from kernel arguments that would occupy scalar registers for the
// whole tracer loop.
#pragma once

#include "kerr_dev.h"
#include "params.h"

namespace bhx {

__device__ __forceinline__ double pow_pos(double x, double y) { return exp_near(y * log_pos(x)); }

__device__ __forceinline__ double smooth01(double t) {
    t = dclamp(t, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

// Quintic step 6t^5 - 15t^4 + 10t^3 on [0, 1]: continuous second derivative.
__device__ __forceinline__ double smooth_c2(double t) {
    t = dclamp(t, 0.0, 1.0);
    return t * t * t * (10.0 + t * (6.0 * t - 15.0));
}

// Jet::shape (jet.h:191-203): dimensionless comoving emissivity at (z, rho).
__device__ inline double jet_shape(const EventParams& J, double z, double rho) {
    const double az = fabs(z);
    if (az <= J.z_base || az >= J.z_top) return 0.0;
    const double lz = log_pos(dmax(az, 1e-6) / J.z_base);
    const double R = J.R_base * exp_near(J.k * lz);
    const double u = rho / R;
    const double s = (u - J.u_peak) / J.u_wid;
    const double t = u / J.spine_wid;
    const double f = exp_near(-s * s) + J.spine * exp_near(-t * t);
    if (f < 1e-9) return 0.0;
    const double ts = dmin(1.0, pow_pos(dmax(az, 1e-6) / J.z_sat, J.k));
    const double gam = 1.0 + (J.gamma_max - 1.0) * ts;
    // (Gamma/Gamma_base)^-(2+alpha) (R/R_base)^-(3+alpha) (z/z_base)^zeta, R/R_base = (z/z_base)^k
    const double g = pow_pos(gam / J.gamma_zb, -(2.0 + J.alpha)) *
                     exp_near((J.zeta - J.k * (3.0 + J.alpha)) * lz);
    const double ramp = smooth01((az - J.z_base) / (0.5 * J.z_base)) *
                        smooth01((J.z_top - az) / (0.15 * J.z_top));
    return f * g * ramp;
}

// g^(2+alpha) of the jet plasma seen along momentum p at (r, theta):
// Jet::four_velocity (jet.h:211-241) on the ZAMO tetrad, redshift_factor
// (render.cpp:31-34), Jet::boost (jet.h:314).  Returns 0 when g <= 0.
__device__ inline double jet_boost(const EventParams& J, double a, double a2, double r, double th,
                                   double pt, double pr, double pth, double pph) {
    double sn, cs;
    sincos_near(th, &sn, &cs);
    if (fabs(sn) < 1e-9) sn = (sn < 0.0) ? -1e-9 : 1e-9;
    const double sn2 = sn * sn;
    const double S = r * r + a2 * cs * cs;
    const double D = r * r - 2.0 * r + a2;
    const double A = (r * r + a2) * (r * r + a2) - a2 * D * sn2;
    const double lapse = sqrt(S * D / A);
    const double omega = 2.0 * a * r / A;
    const double e00 = 1.0 / lapse, e03 = omega / lapse;
    const double e11 = sqrt(D / S), e22 = 1.0 / sqrt(S), e33 = sqrt(S / A) / sn;
    const double sqrt_gpp = sn * sqrt(A / S);
    const double az = fabs(r * cs);
    const double ts = dmin(1.0, pow_pos(dmax(az, 1e-6) / J.z_sat, J.k));
    const double gam = 1.0 + (J.gamma_max - 1.0) * ts;
    const double v = sqrt(dmax(0.0, 1.0 - 1.0 / (gam * gam)));
    double v_phi = (J.Omega_F - omega) * sqrt_gpp / dmax(lapse, 1e-12);
    v_phi = dclamp(v_phi, -0.5 * v, 0.5 * v);
    const double v_pol = sqrt(dmax(0.0, v * v - v_phi * v_phi));
    const double sgn = (cs >= 0.0) ? 1.0 : -1.0;
    const double d_r = sgn * cs, d_th = -sgn * sn;
    const double G = 1.0 / sqrt(dmax(1e-12, 1.0 - v * v));
    const double u0 = G * e00;
    const double u1 = G * (v_pol * d_r * e11);
    const double u2 = G * (v_pol * d_th * e22);
    const double u3 = G * (e03 + v_phi * e33);
    const double ku = pt * u0 + pr * u1 + pth * u2 + pph * u3;
    if (!(ku > 1e-12)) return 0.0;
    return pow_pos(1.0 / ku, 2.0 + J.alpha);
}

// Conical jet (model 1, no counterpart in the reference): comoving emissivity
// profile at (r, cos theta) - (r0 / r)^2 inside two cones of half opening
// cone_angle about the spin axis, tapered in |cos theta| between the cone's
// edge and 0.8 of its opening and ramped at both ends (the end ramps of the
// reference's jets, jet.h:199-202), all with the quintic step, so that the
// profile has a continuous second derivative and its path integral a smooth
// gradient.
__device__ inline double cone_shape(const EventParams& J, double r, double cs) {
    if (r <= J.cone_r0 || r >= J.cone_r1) return 0.0;
    const double mu = fabs(cs);
    if (mu <= J.cone_cos) return 0.0;   // outside the cone
    const double ta = smooth_c2((mu - J.cone_cos) * J.cone_itaper);
    const double q = J.cone_r0 / r;
    const double ramp = smooth_c2((r - J.cone_r0) / (0.5 * J.cone_r0)) *
                        smooth_c2((J.cone_r1 - r) / (0.15 * J.cone_r1));
    return q * q * ta * ramp;
}

// g^(2+alpha) of the conical jet's plasma, radial outflow u = Gamma (e_t +
// beta e_r) on the ZAMO tetrad, seen along momentum p, at (r, theta) with
// (sin, cos) theta = (sn, cs).  0 when g <= 0.
__device__ inline double cone_boost(const EventParams& J, double a, double a2, double r, double sn, double cs,
                                    double pt, double pr, double pph) {
    if (fabs(sn) < 1e-9) sn = (sn < 0.0) ? -1e-9 : 1e-9;
    const double sn2 = sn * sn;
    const double S = r * r + a2 * cs * cs;
    const double D = r * r - 2.0 * r + a2;
    const double A = (r * r + a2) * (r * r + a2) - a2 * D * sn2;
    const double iA = rcp(A);
    const double ilapse = rcp(sqrt(S * D * iA));
    const double omega = 2.0 * a * r * iA;
    // k.u = Gamma ((p_t + omega p_phi) / lapse + beta p_r sqrt(D / S))
    const double ku = J.cone_gamma * ((pt + omega * pph) * ilapse + J.cone_beta * pr * sqrt(D * rcp(S)));
    if (!(ku > 1e-12)) return 0.0;
    return pow_pos(rcp(ku), 2.0 + J.alpha);
}

// bl_to_cartesian (geodesic.h:256-262).
__device__ __forceinline__ void to_cart(double a2, double r, double th, double ph, double o[3]) {
    double sn, cs, sp, cp;
    sincos_near(th, &sn, &cs);
    sincos_near(ph, &sp, &cp);
    const double rr = sqrt(r * r + a2);
    o[0] = rr * sn * cp;
    o[1] = rr * sn * sp;
    o[2] = r * cs;
}

// segment_distance_to_origin (render.cpp:55-62).
__device__ __forceinline__ double seg_dist(const double p0[3], const double p1[3]) {
    const double d0 = p1[0] - p0[0], d1 = p1[1] - p0[1], d2 = p1[2] - p0[2];
    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
    double t = 0.0;
    if (dd > 0.0) t = dclamp(-(p0[0] * d0 + p0[1] * d1 + p0[2] * d2) / dd, 0.0, 1.0);
    const double q0 = p0[0] + t * d0, q1 = p0[1] + t * d1, q2 = p0[2] + t * d2;
    return sqrt(q0 * q0 + q1 * q1 + q2 * q2);
}

}  // namespace bhx
