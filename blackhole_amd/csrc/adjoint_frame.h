// adjoint_frame.h - host set-up of a differentiable render: the per-frame
// constants of adjoint_dev.h and the derivatives of the spin- and
// camera-dependent ones.  Values come from the same host set-up as the
// single-frame path (so the forward image is that path's image); derivatives
// come from the same closed forms evaluated with dual numbers.
#pragma once

#include "adjoint_dev.h"
#include "host_setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bhx {
namespace adj {

// ISCO (Bardeen, Press & Teukolsky 1972, src/kerr.h:222-229) and the specific
// energy of a circular orbit (src/kerr.h:232-237), generic in the scalar type.
template <class T>
T isco_t(const T& a, int sg) {
    const T Z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    const T Z2 = sqrt(3.0 * a * a + Z1 * Z1);
    return 3.0 + Z2 - static_cast<double>(sg) * sqrt((3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2));
}
template <class T>
T circ_energy_t(const T& a, const T& r, int sg) {
    const T sr = sqrt(r);
    return (r * r - 2.0 * r + static_cast<double>(sg) * a * sr) /
           (r * sqrt(r * r - 3.0 * r + 2.0 * static_cast<double>(sg) * a * sr));
}

// Frame constants of scene s with checkpoints every K accepted steps.
// Returns false for scenes this render cannot differentiate.
inline bool make_frame(const bhx_scene& s, int K, Frame& F, FrameCam& cam) {
    if (!s.disc_enabled || s.ns_enabled || s.sun_enabled || s.hs_enabled) return false;
    if (s.jet_enabled && s.jet_model != 1) return false;
    if (K < 1 || s.max_steps < 1) return false;
    std::memset(&F, 0, sizeof F);
    std::memset(&cam, 0, sizeof cam);
    cam.G = host::raygen_args(s);
    const TraceArgs T = host::trace_args(s);
    const ShadeArgs S = host::shade_args(s);
    F.a = T.a; F.a2 = T.a2; F.r_h = T.r_h; F.r_stop = T.r_stop; F.rtol = T.rtol; F.atol = T.atol;
    F.r_escape = T.r_escape; F.r_in = T.r_in; F.r_out = T.r_out;
    cam.h0 = T.h0;
    F.max_steps = T.max_steps;
    F.K = K;
    F.sg = S.disc_sg;
    F.disc_a = S.disc_a;
    F.flux_scale = S.flux_scale;
    F.eddington = s.eddington;
    F.pt_x0 = S.pt_x0;
    for (int i = 0; i < 3; ++i) { F.pt_c[i] = S.pt_c[i]; F.pt_x[i] = S.pt_x[i]; }
    // spin derivatives of the disc's spin-only constants (src/disc.h:46-82)
    {
        using D1 = Dual<1>;
        const D1 a = D1::var(S.disc_a, 0);
        const int sg = S.disc_sg;
        const D1 r_in = isco_t(a, sg);
        const D1 eta = 1.0 - circ_energy_t(a, r_in, sg);
        // flux_scale = eddington L_Edd / (eta c^2) c^2 / r_g^2, L_Edd and r_g spin-free
        F.dflux_da = -S.flux_scale * eta.d[0] / eta.v;
        const D1 ac = acos(a);
        const D1 x[3] = {2.0 * cos(ac / 3.0 - kPi_ / 3.0), 2.0 * cos(ac / 3.0 + kPi_ / 3.0),
                         -2.0 * cos(ac / 3.0)};
        const D1 x0 = sqrt(r_in);
        F.d_x0 = x0.d[0];
        const int jj[3] = {1, 0, 0}, kk[3] = {2, 2, 1};
        for (int i = 0; i < 3; ++i) {
            F.d_x[i] = x[i].d[0];
            if (S.pt_c[i] != 0.0) {
                const D1 xi = x[i], xj = x[jj[i]], xk = x[kk[i]];
                const D1 ci = 3.0 * (xi - a) * (xi - a) / (xi * (xi - xj) * (xi - xk));
                F.d_c[i] = ci.d[0];
            }
        }
    }
    // conical jet: emission j_nu0 x profile x g^(2+alpha) along the ray; the
    // band integral of its power-law spectrum turns the specific intensity at
    // nu0 into intensity (the normalisation of bhx_setup_conical_jet)
    if (s.jet_enabled && s.jet_j_nu0 > 0.0) {
        const EventParams E = host::event_params(s);
        F.has_jet = 1;
        F.j0 = s.jet_j_nu0;
        F.alpha = s.jet_alpha;
        F.r_bound = s.jet_r_bound;
        F.cone_angle = E.cone_angle; F.cone_r0 = E.cone_r0; F.cone_r1 = E.cone_r1;
        F.cone_beta = E.cone_beta; F.cone_gamma = E.cone_gamma; F.cone_cos = E.cone_cos; F.cone_itaper = E.cone_itaper;
        const double nu_0 = 299792458.0 / 550e-9, nu_lo = 1.0e9, nu_hi = 1.0e17;
        const double p = 1.0 - s.jet_alpha;
        const double band = (std::fabs(p) > 1e-9) ? (std::pow(nu_hi, p) - std::pow(nu_lo, p)) / p
                                                  : std::log(nu_hi / nu_lo);
        F.kappa = std::pow(nu_0, s.jet_alpha) * band;
        // j_nu0 ~ Omega_H^2 f(Omega_H^2) Mdot, Mdot ~ eddington / eta(spin)
        using D1 = Dual<1>;
        const double ad = std::clamp(s.spin, -0.9999, 0.9999);
        const D1 a = D1::var(s.spin, 0);
        const D1 a_d = (ad == s.spin) ? D1::var(ad, 0) : D1(ad);   // eta is flat beyond the clamp
        const D1 eta = 1.0 - circ_energy_t(a_d, isco_t(a_d, s.disc_sense), s.disc_sense);
        const D1 rp = 1.0 + sqrt(1.0 - a * a);
        const D1 om = a / (2.0 * rp);
        const D1 x = om * om;
        const D1 f = 1.0 + 1.38 * x - 9.2 * x * x;
        const D1 P = x * f / eta;
        F.dlnj0_da = P.d[0] / P.v;
    }
    // camera frame (src/kerr.h:44-77, 270-293) as a function of (a, r_cam, theta_cam)
    {
        using D3 = Dual<3>;
        const D3 a = D3::var(s.spin, 0), r = D3::var(s.cam_r, 1), th = D3::var(s.cam_theta, 2);
        const D3 sn = sin(th), cs = cos(th);
        const D3 a2 = a * a;
        const D3 sn2 = sn * sn, cs2 = cs * cs;
        const D3 Sg = r * r + a2 * cs2;
        const D3 Dl = r * r - 2.0 * r + a2;
        const D3 A = (r * r + a2) * (r * r + a2) - a2 * Dl * sn2;
        const D3 alpha = sqrt(Sg * Dl / A);
        const D3 omega = 2.0 * a * r / A;
        const D3 e[10] = {1.0 / alpha, omega / alpha, sqrt(Dl / Sg), 1.0 / sqrt(Sg), sqrt(Sg / A) / sn,
                          -(1.0 - 2.0 * r / Sg), -2.0 * a * r * sn2 / Sg, Sg / Dl, Sg, A * sn2 / Sg};
        for (int i = 0; i < 10; ++i)
            for (int m = 0; m < 3; ++m) cam.dcam[i][m] = e[i].d[m];
    }
    return true;
}

}  // namespace adj
}  // namespace bhx
