// adjoint_dev.h - per-ray pieces of the differentiable render.
//
// Forward: the reference's integration of a photon (src/render.cpp:69-260,
// src/geodesic.h:130-194) restricted to what a disc image needs (horizon,
// equatorial disc crossing, escape, step budget), writing a checkpoint of the
// full integrator state (y, h, controller memory) every K accepted steps.
//
// Backward: the discrete adjoint of that computation with the step-size
// sequence frozen.  Each segment of K accepted steps is replayed from its
// checkpoint (the same arithmetic, so the same accept/reject decisions and
// the same states), its states stored, then swept in reverse with hand-written
// vector-Jacobian products of the Dormand-Prince step and of the Hamiltonian
// right-hand side.  The crossing time s* of the terminal step is defined by
// theta(s*) = pi/2 on the dense output; its derivative comes from the
// implicit-function theorem, ds* = -d theta / theta'(s*).  Only (r, theta,
// p_r, p_theta) carry adjoints: t and phi never feed back into the flow or
// into the disc intensity.
//
// All functions are host-callable so the adjoint can be checked against
// finite differences of itself on the CPU.
#pragma once

#include "dense_dev.h"
#include "dp5.h"
#include "dual.h"
#include "kerr_dev.h"
#include "params.h"
#include "raygen_dev.h"
#include "../../include/bhx.h"

namespace bhx {
namespace adj {

using namespace dp5;

constexpr double kPi_ = 3.14159265358979323846;
constexpr double kSigmaSB_ = 5.670374419e-8;
// A checkpoint record: y[6], h, controller memory (beta log err_prev), every
// K accepted steps.  With K = 1 the forward keeps a record of every accepted
// step instead - (r, theta, p_r, p_theta) at its start, the step size it was
// accepted with and its near-jet flag - and the backward reads them without
// replaying.
constexpr int kCk = 8;
constexpr int kStep = 6;

// Per-frame constants of a differentiable render.
struct Frame {
    // integrator (src/render.cpp:71-86)
    double a, a2, r_h, r_stop, rtol, atol, r_escape, r_in, r_out;
    int32_t max_steps, K, sg, pad_k;
    // Novikov-Thorne disc (src/disc.h:46-121) and the spin derivatives of its
    // spin-only constants
    double disc_a, flux_scale, dflux_da, eddington;
    double pt_x0, pt_c[3], pt_x[3], d_x0, d_c[3], d_x[3];
    // conical jet (bhx_scene.jet_model = 1), integrated along the ray: its
    // intensity kappa * J joins the disc's (kappa = nu0^alpha x band integral,
    // J the specific intensity at nu0); d ln j_nu0 / d spin from the
    // Blandford-Znajek power and the disc efficiency
    int32_t has_jet, pad_;
    double j0, alpha, kappa, dlnj0_da, r_bound;
    double cone_angle, cone_r0, cone_r1, cone_beta, cone_gamma, cone_cos, cone_itaper;
};

// The camera of a frame: ray generation (pixel map, jitter, ZAMO tetrad and
// metric at the camera), the integrator's first step h0 = 0.02 max(1, r_cam)
// (src/render.cpp:86) and the derivatives
// d(e00 e03 e11 e22 e33 gl_tt gl_tp gl_rr gl_hh gl_pp)/d(a, r_cam, theta_cam).
// Frames that differ only in their camera share one Frame.
struct FrameCam {
    RaygenArgs G;
    double h0;
    double dcam[10][3];
};

// Integrator state of one ray.
struct Ray {
    double y[6], k1[6], pt, pph, h, lpw, c_prev, jet;   // lpw = beta log(err_prev)
    int nsteps, tries;
    int jet_wait;   // a jet sample of the last step waits in the caller's slot (see fwd_step)
};

__host__ __device__ inline void write_ck(double* ck, const Ray& s) {
    for (int c = 0; c < 6; ++c) ck[c] = s.y[c];
    ck[6] = s.h;
    ck[7] = s.lpw;
}

// The record of one accepted step (K = 1): the state it started from (t and
// phi never feed back), the step size it was accepted with, whether the jet
// was sampled on it.
__host__ __device__ inline void write_step(double* ck, const Ray& s, bool near) {
    ck[0] = s.y[1];
    ck[1] = s.y[2];
    ck[2] = s.y[4];
    ck[3] = s.y[5];
    ck[4] = s.h;
    ck[5] = near ? 1.0 : 0.0;
}
// Position of a record in a store walked from the last record back.
struct CkPos {
    int32_t c, j;   // chunk (stores that have them), record
};

// A ray's checkpoints in one flat array: record j at base + j kCk.  The
// forward and the backward take any store with put(j) (the slot to write
// record j to, or NULL to drop it; records come in order), get(j), and the
// walk back from the last of n records, rpos(n) / at(p) / step_back(p); the
// kernels use a chunked pool (adjoint.hip).
struct FlatCk {
    double* base;
    __host__ __device__ double* put(int j) { return base + static_cast<int64_t>(j) * kCk; }
    __host__ __device__ const double* get(int j) const { return base + static_cast<int64_t>(j) * kCk; }
    __host__ __device__ CkPos rpos(int n) const { return {0, n - 1}; }
    __host__ __device__ const double* at(const CkPos& p) const { return get(p.j); }
    __host__ __device__ void step_back(CkPos& p) const { --p.j; }
};

// One DP5 attempt from s (the tracer's arithmetic, src/geodesic.h:130-169).
// u = err^2 (err the RMS scaled error; 1e20 when not finite): the square root
// is never taken, see accepted() and control_fac().
struct Att {
    double yn[6], c5[6], u;
    Rhs g7;
};
__host__ __device__ inline void attempt(const RayK& K, const Ray& s, double rtol, double atol, Att& o) {
    const double h = s.h;
    const double* y = s.y;
    const double* k1 = s.k1;
    double tmp[6], k2[6], k3[6], k4[6], k5[6], k6[6];
    Rhs f;
    for (int c = 0; c < 6; ++c) tmp[c] = y[c] + h * A21 * k1[c];
    kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
    for (int c = 0; c < 6; ++c) { k2[c] = f.d[c]; tmp[c] = y[c] + h * (A31 * k1[c] + A32 * k2[c]); }
    kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
    for (int c = 0; c < 6; ++c) {
        k3[c] = f.d[c];
        tmp[c] = y[c] + h * (A41 * k1[c] + A42 * k2[c] + A43 * k3[c]);
    }
    kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
    for (int c = 0; c < 6; ++c) {
        k4[c] = f.d[c];
        tmp[c] = y[c] + h * (A51 * k1[c] + A52 * k2[c] + A53 * k3[c] + A54 * k4[c]);
    }
    kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
    for (int c = 0; c < 6; ++c) {
        k5[c] = f.d[c];
        tmp[c] = y[c] + h * (A61 * k1[c] + A62 * k2[c] + A63 * k3[c] + A64 * k4[c] + A65 * k5[c]);
    }
    kerr_rhs(K, tmp[1], tmp[2], tmp[4], tmp[5], f);
    for (int c = 0; c < 6; ++c) {
        k6[c] = f.d[c];
        o.yn[c] = y[c] + h * (B1 * k1[c] + B3 * k3[c] + B4 * k4[c] + B5 * k5[c] + B6 * k6[c]);
    }
    kerr_rhs(K, o.yn[1], o.yn[2], o.yn[4], o.yn[5], o.g7);
    double err2 = 0.0;
    for (int c = 0; c < 6; ++c) {
        const double e = h * (E1 * k1[c] + E3 * k3[c] + E4 * k4[c] + E5 * k5[c] + E6 * k6[c] + E7 * o.g7.d[c]);
        const double sc = atol + rtol * fmax(fabs(y[c]), fabs(o.yn[c]));   // y is never NaN
        const double qe = e * rcp_coarse(sc);
        err2 += qe * qe;
        o.c5[c] = h * (D1 * k1[c] + D3 * k3[c] + D4 * k4[c] + D5 * k5[c] + D6 * k6[c] + D7 * o.g7.d[c]);
    }
    // err = 1e10 when not finite (geodesic.h:166): min(u, 1e20) maps inf and
    // NaN there, and any u >= 1e20 is rejected with fac clamped to 0.2 alike
    o.u = fmin(err2 * 0.125, 1e20);
}

// err <= 1 (geodesic.h:175): sqrt(u) <= 1 exactly when u <= 1 + 2^-52.
__host__ __device__ inline bool accepted(const Att& A) { return A.u <= 1.0 + 0x1p-52; }

// hmax clamp before a step (src/render.cpp:130-131), once per step.
__host__ __device__ inline void clamp_h(const Frame& F, Ray& s) {
    const double hmax = 0.05 * dmax(1.0, s.y[1] - F.r_h) + 0.01;
    if (s.tries == 0 && s.h > hmax) s.h = hmax;
}

// PI controller factor (src/geodesic.h:170-173), 0.9 err^-alpha
// err_prev^beta as one exponential, and the log of the error (log(u) / 2).
__host__ __device__ inline double control_fac(const Ray& s, double u, double* lg_out) {
    const double lg = 0.5 * log_pos(fmax(u, 1e-16 * 1e-16));
    *lg_out = lg;
    return fmin(fmax(0.9 * exp_near(s.lpw - kAlpha * lg), 0.2), 5.0);   // never NaN: std::clamp
}

// Accepted step that does not end the ray: advance the state.
__host__ __device__ inline void advance(Ray& s, const Att& A, double fac, double lg) {
    for (int c = 0; c < 6; ++c) { s.y[c] = A.yn[c]; s.k1[c] = A.g7.d[c]; }
    s.c_prev = A.g7.cs;
    s.lpw = kBeta * dmax(lg, kLog1em4);   // err_prev = max(err, 1e-4)
    s.h *= fac;
}

// Dense-output coefficients of one component (src/geodesic.h:177-185).
__host__ __device__ inline void dense_coeffs(const Ray& s, const Att& A, int c, double o[5]) {
    o[0] = s.y[c];
    o[1] = A.yn[c] - s.y[c];
    o[2] = s.h * s.k1[c] - o[1];
    o[3] = o[1] - s.h * A.g7.d[c] - o[2];
    o[4] = A.c5[c];
}

// ---- disc emission (src/disc.h:94-121, src/render.cpp:31-34, 280-300) ------

// Bolometric intensity sigma (g T)^4 / pi of a disc hit at r seen along a ray
// with conserved p_t, p_phi, exactly as the shading computes T and g.
__host__ __device__ inline double disc_intensity(const Frame& F, double r, double pt, double pph) {
    if (r <= F.r_in || r > F.r_out) return 0.0;
    const double a = F.disc_a, sg = (double)F.sg;
    const double sr = sqrt(r);
    const double r32 = r * sr;
    const double Om = sg / (r32 + sg * a);
    const double dd = r32 + sg * a;
    const double dOm = -sg * 1.5 * sr / (dd * dd);
    const double root = sqrt(r * r - 3.0 * r + 2.0 * sg * a * sr);
    const double E = (r * r - 2.0 * r + sg * a * sr) / (r * root);
    const double L = (sg * sr * (r * r - 2.0 * sg * a * sr + a * a)) / (r * root);
    const double den = E - Om * L;
    if (!(den > 0.0)) return 0.0;
    const double x = sqrt(r);
    double W = x - F.pt_x0 - 1.5 * a * log(x / F.pt_x0);
    for (int i = 0; i < 3; ++i)
        if (F.pt_c[i] != 0.0) W -= F.pt_c[i] * log((x - F.pt_x[i]) / (F.pt_x0 - F.pt_x[i]));
    double Fl = (-dOm) / (4.0 * kPi_ * r * den * den) * W;
    if (!(Fl > 0.0 && finite_d(Fl))) return 0.0;
    Fl = Fl * F.flux_scale;
    if (Fl <= 0.0) return 0.0;
    const double T = pow(Fl / kSigmaSB_, 0.25);
    const double dsc = r * r * r - 3.0 * r * r + 2.0 * sg * a * r32;
    const double ut = (r32 + sg * a) / sqrt(dmax(1e-300, dsc));
    const double ku = pt * ut + pph * (Om * ut);
    const double g = (ku > 1e-12) ? 1.0 / ku : 0.0;
    if (!(g > 0.0)) return 0.0;
    const double gT = g * T;
    return kSigmaSB_ * (gT * gT) * (gT * gT) / kPi_;
}

// The same intensity as g^4 F / pi with its partials in (r, a, p_t, p_phi);
// a enters through the disc's spin-only constants too (their derivatives are
// in the frame).
__host__ __device__ inline void disc_intensity_grad(const Frame& F, double r, double pt, double pph, double dI[4]) {
    using D = Dual<4>;
    const D R = D::var(r, 0), A = D::var(F.disc_a, 1), PT = D::var(pt, 2), PPH = D::var(pph, 3);
    const double sg = (double)F.sg;
    auto cst = [](double v, double da) {
        D c(v);
        c.d[1] = da;
        return c;
    };
    const D sr = sqrt(R);
    const D r32 = R * sr;
    const D dd = r32 + sg * A;
    const D Om = sg / dd;
    const D dOm = -(sg * 1.5) * sr / (dd * dd);
    const D root = sqrt(R * R - 3.0 * R + 2.0 * sg * A * sr);
    const D E = (R * R - 2.0 * R + sg * A * sr) / (R * root);
    const D L = (sg * sr * (R * R - 2.0 * sg * A * sr + A * A)) / (R * root);
    const D den = E - Om * L;
    const D x0 = cst(F.pt_x0, F.d_x0);
    D W = sr - x0 - 1.5 * A * log(sr / x0);
    for (int i = 0; i < 3; ++i)
        if (F.pt_c[i] != 0.0) {
            const D ci = cst(F.pt_c[i], F.d_c[i]), xi = cst(F.pt_x[i], F.d_x[i]);
            W = W - ci * log((sr - xi) / (x0 - xi));
        }
    const D Fl = (-dOm) / (4.0 * kPi_ * R * den * den) * W * cst(F.flux_scale, F.dflux_da);
    const D dsc = R * R * R - 3.0 * R * R + 2.0 * sg * A * r32;
    const D ut = (r32 + sg * A) / sqrt(dsc);
    const D g = 1.0 / (PT * ut + PPH * (Om * ut));
    const D g2 = g * g;
    const D I = g2 * g2 * Fl / kPi_;
    for (int k = 0; k < 4; ++k) dI[k] = I.d[k];
}

// ---- conical jet -------------------------------------------------------------

template <class T>
__host__ __device__ inline T smooth_c2_t(const T& t) {
    if (value(t) <= 0.0) return T(0.0);
    if (value(t) >= 1.0) return T(1.0);
    return t * t * t * (10.0 + t * (6.0 * t - 15.0));
}
template <class T>
__host__ __device__ inline T abs_t(const T& x) { return (value(x) < 0.0) ? -x : x; }
// f(x) from its value v and derivative dv at value(x) (sin and cos from one sincos)
__host__ __device__ inline double with_value(double, double v, double) { return v; }
template <int N>
__host__ __device__ inline Dual<N> with_value(const Dual<N>& x, double v, double dv) { return chain(x, v, dv); }

// Integrand of the conical jet at dense-output state (r, theta, p_r) of a
// ray with constants (a, p_t, p_phi): comoving profile (r0/r)^2 x taper x
// ramps, times g^(2+alpha) of the radial Gamma outflow on the ZAMO tetrad -
// the device's cone_shape x cone_boost, generic in the scalar type so that
// dual numbers give its gradient.
template <class T>
__host__ __device__ inline T cone_F(const Frame& F, const T& r, const T& th, const T& pr, const T& a, const T& pt,
                                    const T& pph) {
    if (value(r) <= F.cone_r0 || value(r) >= F.cone_r1) return T(0.0);
    double snv, csv;
    sincos_near(value(th), &snv, &csv);
    if (fabs(csv) <= F.cone_cos) return T(0.0);   // outside the cone: the taper is zero
    T sn = with_value(th, snv, csv);
    const T cs = with_value(th, csv, -snv);
    const T ta = smooth_c2_t((abs_t(cs) - F.cone_cos) * F.cone_itaper);
    const T q = F.cone_r0 / r;
    const T ramp = smooth_c2_t((r - F.cone_r0) / (0.5 * F.cone_r0)) * smooth_c2_t((F.cone_r1 - r) / (0.15 * F.cone_r1));
    if (fabs(value(sn)) < 1e-9) sn = T((value(sn) < 0.0) ? -1e-9 : 1e-9);
    const T a2 = a * a;
    const T S = r * r + a2 * cs * cs;
    const T D = r * r - 2.0 * r + a2;
    const T A = (r * r + a2) * (r * r + a2) - a2 * D * sn * sn;
    const T lapse = sqrt(S * D / A);
    const T omega = 2.0 * a * r / A;
    const T ku = pt * (F.cone_gamma / lapse) + pr * (F.cone_gamma * F.cone_beta * sqrt(D / S)) +
                 pph * (F.cone_gamma * omega / lapse);
    if (!(value(ku) > 1e-12)) return T(0.0);
    return q * q * ta * ramp * pow(1.0 / ku, 2.0 + F.alpha);
}

// pseudo-Cartesian point (geodesic.h:256-262), the segment's distance to
// the origin (render.cpp:55-62) and the conical jet's sampling gate: the
// jet's cull, as the tracer does it
__host__ __device__ inline void cart_hd(double a2, double r, double th, double ph, double o[3]) {
    double sn, cs, sp, cp;
    sincos_near(th, &sn, &cs);
    sincos_near(ph, &sp, &cp);
    const double rr = sqrt(r * r + a2);
    o[0] = rr * sn * cp;
    o[1] = rr * sn * sp;
    o[2] = r * cs;
}
__host__ __device__ inline bool near_jet(const Frame& F, const Ray& s, const Att& A) {
    if (!F.has_jet) return false;
    const double* y = s.y;
    const double* yn = A.yn;
    double p0[3], p1[3];
    cart_hd(F.a2, y[1], y[2], y[3], p0);
    cart_hd(F.a2, yn[1], yn[2], yn[3], p1);
    const double d0 = p1[0] - p0[0], d1 = p1[1] - p0[1], d2 = p1[2] - p0[2];
    const double dd = d0 * d0 + d1 * d1 + d2 * d2;
    double t = 0.0;
    if (dd > 0.0) t = dclamp(-(p0[0] * d0 + p0[1] * d1 + p0[2] * d2) / dd, 0.0, 1.0);
    const double q0 = p0[0] + t * d0, q1 = p0[1] + t * d1, q2 = p0[2] + t * d2;
    if (!(sqrt(q0 * q0 + q1 * q1 + q2 * q2) < F.r_bound)) return false;
    // the conical jet's gate: the step's dense output can reach the cone
    const double c2t = yn[2] - y[2];
    const double c3t = s.h * s.k1[2] - c2t;
    return cone_reach(y[2], c2t, c3t, c2t - s.h * A.g7.d[2] - c3t, A.c5[2], F.cone_angle);
}

// Four-point Gauss-Legendre nodes and weights on [0, 1] (render.cpp:47-50).
__host__ __device__ inline double gl_x(int q) {
    return q == 0 ? 0.06943184420297371 : q == 1 ? 0.33000947820757187 : q == 2 ? 0.66999052179242813
                                                                              : 0.93056815579702629;
}
__host__ __device__ inline double gl_w(int q) { return (q == 0 || q == 3) ? 0.17392742256872693 : 0.32607257743127307; }

// The jet's contribution of one step over the fraction [0, s_end] of its
// dense output cf (components r, theta, p_r, p_theta): kConeSub
// sub-intervals of 4-point Gauss-Legendre, j0 h s_end sum w_q F / kConeSub
// (the tracer's sampling of the conical jet).
// The integrand at node j (sub-interval j / 4, Gauss point j % 4) of a step's
// jet sample over [0, s_end] of the dense output c15 (r, theta, p_r: 5
// coefficients each).
__host__ __device__ inline double jet_value_node(const Frame& F, const double* c15, double s_end, double pt,
                                                 double pph, int j) {
    const double u = ((j >> 2) + gl_x(j & 3)) / kConeSub;
    const double sv = u * s_end;
    const double r = dense_comp(c15[0], c15[1], c15[2], c15[3], c15[4], sv);
    const double th = dense_comp(c15[5], c15[6], c15[7], c15[8], c15[9], sv);
    const double pr = dense_comp(c15[10], c15[11], c15[12], c15[13], c15[14], sv);
    return cone_F<double>(F, r, th, pr, F.a, pt, pph);
}

__host__ __device__ inline double jet_step(const Frame& F, const double cf[4][5], double h, double s_end,
                                           double pt, double pph) {
    double c15[15];
    for (int c = 0; c < 3; ++c)
        for (int m = 0; m < 5; ++m) c15[5 * c + m] = cf[c][m];
    double sum = 0.0;
#pragma unroll 1
    for (int j = 0; j < 4 * kConeSub; ++j) sum += gl_w(j & 3) * jet_value_node(F, c15, s_end, pt, pph, j);
    return F.j0 * sum * h * s_end / kConeSub;
}

// ---- reverse mode of the right-hand side -----------------------------------

// Vector-Jacobian product of the (r, theta, p_r, p_theta) rates of kerr_rhs
// (the separable Hamiltonian flow, src/geodesic.h:50-71 with the metric of
// src/kerr.h:107-176) at (r, th, pr, pth): given weights w of (dr, dth, dpr,
// dpth), adds w^T J to xb (w.r.t. r, th, pr, pth) and to cb (w.r.t. a, p_t,
// p_phi).  Written from the forward expressions line by line in reverse.
__host__ __device__ inline void rhs_vjp(const RayK& K, const double x[4], const double w[4], double xb[4],
                                        double cb[3]) {
    const double r = x[0], th = x[1], pr = x[2], pth = x[3];
    const double a = K.a, a2 = K.a2, pt = K.pt, pph = K.pph;
    double s, c;
    sincos_near(th, &s, &c);
    if (fabs(s) < 1e-9) s = (s < 0.0) ? -1e-9 : 1e-9;
    const double s2 = s * s, c2 = c * c;
    const double r2 = r * r;
    const double rpa = r2 + a2;
    const double S = r2 + a2 * c2;
    const double D = rpa - 2.0 * r;
    // one reciprocal of Sigma Delta sin^2 for the three (as kerr_rhs_sc forms
    // them; all three factors are positive on a ray outside the horizon)
    const double SD = S * D;
    const double iX = rcp(SD * s2);
    const double iS = D * s2 * iX, iD = S * s2 * iX, is2 = SD * iX;
    const double P = rpa * pt + a * pph;
    const double PiD = P * iD;
    const double U = pph + a * pt * s2;
    const double pr2 = pr * pr;
    const double N = D * pr2 + pth * pth + U * U * is2 - P * PiD;
    const double H = N * iS;
    const double Dr = 2.0 * r - 2.0;
    const double br = Dr * pr2 + PiD * (Dr * PiD - 4.0 * r * pt) - 2.0 * r * H;
    const double bt = a2 * pt * pt + a2 * H - pph * pph * is2 * is2;
    const double q = s * c;
    const double w1 = w[0], w2 = w[1], w4 = w[2], w5 = w[3];
    double rb = 0.0, thb = 0.0, prb = 0.0, pthb = 0.0, ab = 0.0, ptb = 0.0, pphb = 0.0;
    double Sb = 0.0, Db = 0.0, iSb = 0.0, iDb = 0.0, is2b = 0.0, Pb = 0.0, Ub = 0.0, Nb = 0.0, Hb = 0.0;
    double Drb = 0.0, brb = 0.0, btb = 0.0, qb = 0.0, s2b = 0.0, c2b = 0.0, a2b = 0.0, PiDb = 0.0, rpab = 0.0;
    // d[1] = D pr iS
    Db += pr * iS * w1; prb += D * iS * w1; iSb += D * pr * w1;
    // d[2] = pth iS
    pthb += iS * w2; iSb += pth * w2;
    // d[4] = -1/2 iS br
    iSb += -0.5 * br * w4; brb += -0.5 * iS * w4;
    // d[5] = -(s c) iS bt
    qb += -iS * bt * w5; iSb += -q * bt * w5; btb += -q * iS * w5;
    // bt = a^2 pt^2 + a^2 H - pph^2 is2^2
    a2b += (pt * pt + H) * btb; ptb += 2.0 * a2 * pt * btb; Hb += a2 * btb;
    pphb += -2.0 * pph * is2 * is2 * btb; is2b += -2.0 * pph * pph * is2 * btb;
    // br = Dr pr^2 + Dr PiD^2 - 4 r pt PiD - 2 r H
    Drb += (pr2 + PiD * PiD) * brb;
    prb += 2.0 * Dr * pr * brb;
    PiDb += (2.0 * Dr * PiD - 4.0 * r * pt) * brb;
    rb += (-4.0 * pt * PiD - 2.0 * H) * brb;
    ptb += -4.0 * r * PiD * brb;
    Hb += -2.0 * r * brb;
    rb += 2.0 * Drb;                                     // Dr = 2r - 2
    Nb += iS * Hb; iSb += N * Hb;                        // H = N iS
    // N = D pr^2 + pth^2 + U^2 is2 - P PiD
    Db += pr2 * Nb; prb += 2.0 * D * pr * Nb; pthb += 2.0 * pth * Nb;
    Ub += 2.0 * U * is2 * Nb; is2b += U * U * Nb; Pb += -PiD * Nb; PiDb += -P * Nb;
    Pb += iD * PiDb; iDb += P * PiDb;                    // PiD = P iD
    pphb += Ub; ab += pt * s2 * Ub; ptb += a * s2 * Ub; s2b += a * pt * Ub;   // U = pph + a pt s^2
    rpab += pt * Pb; ptb += rpa * Pb; ab += pph * Pb; pphb += a * Pb;        // P = rpa pt + a pph
    Sb += -iS * iS * iSb; Db += -iD * iD * iDb; s2b += -is2 * is2 * is2b;    // reciprocals
    rpab += Db; rb += -2.0 * Db;                         // D = rpa - 2r
    rb += 2.0 * r * Sb; a2b += c2 * Sb; c2b += a2 * Sb;  // S = r^2 + a^2 c^2
    rb += 2.0 * r * rpab; a2b += rpab;                   // rpa = r^2 + a^2
    const double sbar = 2.0 * s * s2b + c * qb, cbar = 2.0 * c * c2b + s * qb;
    thb += c * sbar - s * cbar;
    ab += 2.0 * a * a2b;
    xb[0] += rb; xb[1] += thb; xb[2] += prb; xb[3] += pthb;
    cb[0] += ab; cb[1] += ptb; cb[2] += pphb;
}

// (dr, dth, dpr, dpth) of kerr_rhs at x.
__host__ __device__ inline void rhs_x(const RayK& K, const double x[4], double k[4]) {
    Rhs f;
    kerr_rhs(K, x[0], x[1], x[2], x[3], f);
    k[0] = f.d[1]; k[1] = f.d[2]; k[2] = f.d[4]; k[3] = f.d[5];
}

// Reverse of one Dormand-Prince step x -> yn = x + h sum b_i k_i with h
// frozen.  ynb: adjoint of yn; kb[0..6]: extra adjoints of k1..k7 (k7 =
// f(yn), only with use_k7); adds the adjoint of x to xb and of (a, p_t,
// p_phi) to cb.
__host__ __device__ inline void step_vjp(const RayK& K, const double x[4], double h, double ynb[4],
                                         double kb[7][4], bool use_k7, double xb[4], double cb[3]) {
    const double A[6][5] = {{0, 0, 0, 0, 0},
                            {A21, 0, 0, 0, 0},
                            {A31, A32, 0, 0, 0},
                            {A41, A42, A43, 0, 0},
                            {A51, A52, A53, A54, 0},
                            {A61, A62, A63, A64, A65}};
    const double B[6] = {B1, 0.0, B3, B4, B5, B6};
    double X[6][4], k[6][4];
    for (int i = 0; i < 6; ++i) {
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
            for (int j = 0; j < i; ++j) acc += A[i][j] * k[j][c];
            X[i][c] = (i == 0) ? x[c] : x[c] + h * acc;
        }
        rhs_x(K, X[i], k[i]);
    }
    if (use_k7) {
        double yn[4];
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
            for (int j = 0; j < 6; ++j) acc += B[j] * k[j][c];
            yn[c] = x[c] + h * acc;
        }
        rhs_vjp(K, yn, kb[6], ynb, cb);
    }
    for (int c = 0; c < 4; ++c) {
        xb[c] += ynb[c];
        for (int j = 0; j < 6; ++j) kb[j][c] += h * B[j] * ynb[c];
    }
    for (int i = 5; i >= 0; --i) {
        double Xb[4] = {0.0, 0.0, 0.0, 0.0};
        rhs_vjp(K, X[i], kb[i], Xb, cb);
        for (int c = 0; c < 4; ++c) {
            xb[c] += Xb[c];
            for (int j = 0; j < i; ++j) kb[j][c] += h * A[i][j] * Xb[c];
        }
    }
}

// Dense-output coefficients (c1..c5, Hairer's form, geodesic.h:177-185) of
// the four components of the frozen step from x with size h.
__host__ __device__ inline void step_dense(const RayK& K, const double x[4], double h, double cf[4][5]) {
    const double A[6][5] = {{0, 0, 0, 0, 0},
                            {A21, 0, 0, 0, 0},
                            {A31, A32, 0, 0, 0},
                            {A41, A42, A43, 0, 0},
                            {A51, A52, A53, A54, 0},
                            {A61, A62, A63, A64, A65}};
    const double B[6] = {B1, 0.0, B3, B4, B5, B6};
    const double Dd[7] = {D1, 0.0, D3, D4, D5, D6, D7};
    double k[7][4], X[4], yn[4];
    for (int i = 0; i < 6; ++i) {
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
            for (int j = 0; j < i; ++j) acc += A[i][j] * k[j][c];
            X[c] = (i == 0) ? x[c] : x[c] + h * acc;
        }
        rhs_x(K, X, k[i]);
    }
    for (int c = 0; c < 4; ++c) {
        double acc = 0.0;
        for (int j = 0; j < 6; ++j) acc += B[j] * k[j][c];
        yn[c] = x[c] + h * acc;
    }
    rhs_x(K, yn, k[6]);
    for (int c = 0; c < 4; ++c) {
        cf[c][0] = x[c];
        cf[c][1] = yn[c] - x[c];
        cf[c][2] = h * k[0][c] - cf[c][1];
        cf[c][3] = cf[c][1] - h * k[6][c] - cf[c][2];
        double acc = 0.0;
        for (int j = 0; j < 7; ++j) acc += Dd[j] * k[j][c];
        cf[c][4] = h * acc;
    }
}

// A seed Yb on the dense output at s, as adjoints of the coefficients.
__host__ __device__ inline void dense_seed(double s, const double Yb[4], double cbar[5][4]) {
    const double s1 = 1.0 - s;
    const double wts[5] = {1.0, s, s * s1, s * s * s1, s * s * s1 * s1};
    for (int m = 0; m < 5; ++m)
        for (int c = 0; c < 4; ++c) cbar[m][c] += wts[m] * Yb[c];
}

// Reverse of the coefficients' definitions: adjoints of c1..c5 become
// adjoints of x, of yn and of the stage slopes k1..k7.
__host__ __device__ inline void dense_coef_vjp(double h, const double cbar[5][4], double xb[4], double ynb[4],
                                               double kb[7][4]) {
    const double Dd[7] = {D1, 0.0, D3, D4, D5, D6, D7};
    for (int c = 0; c < 4; ++c) {
        double c2b = cbar[1][c], c3b = cbar[2][c];
        const double c4b = cbar[3][c], c5b = cbar[4][c];
        for (int j = 0; j < 7; ++j) kb[j][c] += h * Dd[j] * c5b;      // c5 = h sum d_j k_j
        c2b += c4b; kb[6][c] += -h * c4b; c3b += -c4b;                 // c4 = c2 - h k7 - c3
        kb[0][c] += h * c3b; c2b += -c3b;                              // c3 = h k1 - c2
        ynb[c] += c2b; xb[c] += -c2b;                                  // c2 = yn - x
        xb[c] += cbar[0][c];                                           // c1 = x
    }
}

// Node j (sub-interval j / 4, Gauss point j % 4) of a step's jet sample over
// [0, s_end] of the dense output cf3 (components r, theta, p_r): the
// integrand and its partials in (r, theta, p_r, a, p_t, p_phi), and the
// s-derivatives of (r, theta, p_r) there.
struct JetNode {
    double f[7];    // value, d/d(r, theta, p_r, a, p_t, p_phi)
    double yp[3];   // d(r, theta, p_r)/ds
};
__host__ __device__ inline void jet_node(const Frame& F, const double cf3[3][5], double s_end, double pt, double pph,
                                         int j, JetNode& o) {
    using D6 = Dual<6>;
    const int k = j >> 2, q = j & 3;
    const double u = (k + gl_x(q)) / kConeSub;
    const double sv = u * s_end;
    double Y[3];
    for (int c = 0; c < 3; ++c) dense_vd(cf3[c][0], cf3[c][1], cf3[c][2], cf3[c][3], cf3[c][4], sv, Y[c], o.yp[c]);
    const D6 f = cone_F<D6>(F, D6::var(Y[0], 0), D6::var(Y[1], 1), D6::var(Y[2], 2), D6::var(F.a, 3),
                            D6::var(pt, 4), D6::var(pph, 5));
    o.f[0] = f.v;
    for (int m = 0; m < 6; ++m) o.f[m + 1] = f.d[m];
}

// Adds node j's share of the seeds of wk J_step (coef = wk j0 h): the node
// seeds on the dense-output coefficients (cbar), the constants' (cb), and its
// terms of J_step's node sum and of d(wk J_step)/ds_end.
__host__ __device__ inline void jet_accum(double coef, double s_end, int j, const JetNode& n, double cbar[5][4],
                                          double cb[3], double* sum, double* dJds) {
    if (n.f[0] == 0.0) return;
    const int k = j >> 2, q = j & 3;
    const double u = (k + gl_x(q)) / kConeSub;
    const double sv = u * s_end;
    *sum += gl_w(q) * n.f[0];
    const double wq = coef * s_end * gl_w(q) / kConeSub;
    const double Yb[4] = {wq * n.f[1], wq * n.f[2], wq * n.f[3], 0.0};
    dense_seed(sv, Yb, cbar);
    cb[0] += wq * n.f[4];
    cb[1] += wq * n.f[5];
    cb[2] += wq * n.f[6];
    *dJds += gl_w(q) / kConeSub * (n.f[0] + s_end * u * (n.f[1] * n.yp[0] + n.f[2] * n.yp[1] + n.f[3] * n.yp[2]));
}

// Adjoint seeds of wk * J_step, J_step = j0 h s_end sum_q w_q F(Y(x_q s_end)),
// on the step's dense output: node seeds into cbar, constants (a, p_t, p_phi)
// into cb.  Returns wk dJ_step/ds_end (for an event-truncated step) and
// J_step in *Js.  The nodes in order, one lane.
__host__ __device__ inline double jet_seeds(const Frame& F, const double cf[4][5], double h, double s_end,
                                            double wk, double pt, double pph, double cbar[5][4], double cb[3],
                                            double* Js) {
    const double coef = wk * F.j0 * h;
    const double cf3[3][5] = {{cf[0][0], cf[0][1], cf[0][2], cf[0][3], cf[0][4]},
                              {cf[1][0], cf[1][1], cf[1][2], cf[1][3], cf[1][4]},
                              {cf[2][0], cf[2][1], cf[2][2], cf[2][3], cf[2][4]}};
    double dJds = 0.0, sum = 0.0;
#pragma unroll 1
    for (int j = 0; j < 4 * kConeSub; ++j) {
        JetNode n;
        jet_node(F, cf3, s_end, pt, pph, j, n);
        jet_accum(coef, s_end, j, n, cbar, cb, &sum, &dJds);
    }
    *Js = F.j0 * sum * h * s_end / kConeSub;
    return coef * dJds;
}

// LDS doubles per wave of jet_seeds_wave: two owners' dense output and
// constants, and their nodes' partials.
constexpr int kJetLds = 2 * 18 + 2 * 4 * kConeSub * 10;
constexpr int kJetShareMax = 8;   // share the nodes when at most this many lanes sample
#if defined(__HIP_DEVICE_COMPILE__)
// jet_seeds with the wave's active lanes sharing the node evaluations: each
// round takes up to two lanes that need a sample (owners), publishes their
// dense output in LDS, spreads their 4 kConeSub nodes over the active lanes,
// and each owner then accumulates its nodes in order - the same arithmetic
// as jet_seeds, so the seeds are the same bits.  All active lanes must call
// it; `need` marks the owners, whose other arguments are used.  lds: this
// wave's kJetLds doubles.
__device__ inline double jet_seeds_wave(const Frame& F, const double cf[4][5], double h, double s_end, double wk,
                                        double pt, double pph, double cbar[5][4], double cb[3], double* Js,
                                        bool need, double* lds) {
    const unsigned long long act = __ballot(true);
    unsigned long long rem = __ballot(need);
    const int n_act = __popcll(act);
    const int lane = __lane_id();
    const int rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(act >> 32),
                                               __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(act), 0u));
    const int per = n_act >= 64 ? 2 : 1;   // owners per round
    double* own_lds = lds;                  // [2][18]: cf3 (15), s_end, pt, pph
    double* node_lds = lds + 2 * 18;        // [2][4 kConeSub][10]: f (7), yp (3)
    double Os = 0.0;
    while (rem != 0ull) {
        int own[2] = {-1, -1};
        own[0] = __builtin_ctzll(rem);
        rem &= rem - 1;
        if (per == 2 && rem != 0ull) {
            own[1] = __builtin_ctzll(rem);
            rem &= rem - 1;
        }
        const int n_own = own[1] >= 0 ? 2 : 1;
        const int oi_self = (lane == own[0]) ? 0 : (lane == own[1]) ? 1 : -1;
        if (oi_self >= 0) {
            double* o = own_lds + 18 * oi_self;
            for (int c = 0; c < 3; ++c)
                for (int m = 0; m < 5; ++m) o[5 * c + m] = cf[c][m];
            o[15] = s_end;
            o[16] = pt;
            o[17] = pph;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int w = rank; w < 4 * kConeSub * n_own; w += n_act) {
            const int oi = w / (4 * kConeSub), j = w - oi * (4 * kConeSub);
            const double* o = own_lds + 18 * oi;
            double cf3[3][5];
            for (int c = 0; c < 3; ++c)
                for (int m = 0; m < 5; ++m) cf3[c][m] = o[5 * c + m];
            JetNode n;
            jet_node(F, cf3, o[15], o[16], o[17], j, n);
            double* d = node_lds + 10 * w;
            for (int m = 0; m < 7; ++m) d[m] = n.f[m];
            for (int m = 0; m < 3; ++m) d[7 + m] = n.yp[m];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (oi_self >= 0) {
            const double coef = wk * F.j0 * h;
            double dJds = 0.0, sum = 0.0;
            for (int j = 0; j < 4 * kConeSub; ++j) {
                const double* d = node_lds + 10 * (oi_self * 4 * kConeSub + j);
                JetNode n;
                for (int m = 0; m < 7; ++m) n.f[m] = d[m];
                for (int m = 0; m < 3; ++m) n.yp[m] = d[7 + m];
                jet_accum(coef, s_end, j, n, cbar, cb, &sum, &dJds);
            }
            *Js = F.j0 * sum * h * s_end / kConeSub;
            Os = coef * dJds;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    return Os;
}
#endif

// Reverse of one frozen step from x with size h: the adjoint xb of its end
// state (ignored for the terminal step, whose end state is not used) becomes
// the adjoint of its start; adjoints of (a, p_t, p_phi) accumulate in cb.
// A step near the jet adds the seeds of wk * its jet sample (over [0, se]
// for the terminal step, else the whole step) and its sample to *J.  The
// terminal (disc-crossing) step seeds Or on r(s*), with s* fixed by
// theta(s*) = pi/2 on the dense output (implicit-function theorem,
// ds* = -d theta(s*) / theta'(s*)).  One routine for every kind of step, so
// the kernel holds one copy of the reverse Dormand-Prince step.
__host__ __device__ inline void step_reverse(const Frame& F, const RayK& K, const double x[4], double h,
                                             bool jet, bool term, double se, double Or, double wk, double xb[4],
                                             double cb[3], double* J, double* lds = nullptr) {
    double ynb[4], kb[7][4] = {};
    for (int c = 0; c < 4; ++c) {
        ynb[c] = term ? 0.0 : xb[c];
        xb[c] = 0.0;
    }
    const bool need = jet && wk != 0.0;
    const bool dense = term || need;
    double cf[4][5], cbar[5][4] = {};
    if (dense) step_dense(K, x, h, cf);
    double Os = 0.0;
#if defined(__HIP_DEVICE_COMPILE__)
    // a few sampling lanes in the wave: every active lane takes part and the
    // nodes are shared; many: each evaluates its own (the same bits either way)
    if (lds && __popcll(__ballot(need)) <= kJetShareMax) {
        double Js = 0.0;
        Os = jet_seeds_wave(F, cf, h, term ? se : 1.0, wk, K.pt, K.pph, cbar, cb, &Js, need, lds);
        if (need) *J += Js;
    } else
#endif
    if (need) {
        double Js = 0.0;
        Os = jet_seeds(F, cf, h, term ? se : 1.0, wk, K.pt, K.pph, cbar, cb, &Js);
        *J += Js;
    }
    if (dense) {
        if (term) {
            double rv, rd, tv, td;
            dense_vd(cf[0][0], cf[0][1], cf[0][2], cf[0][3], cf[0][4], se, rv, rd);
            dense_vd(cf[1][0], cf[1][1], cf[1][2], cf[1][3], cf[1][4], se, tv, td);
            const double Yb[4] = {Or, -(Or * rd + Os) / td, 0.0, 0.0};
            dense_seed(se, Yb, cbar);
        }
        dense_coef_vjp(h, cbar, xb, ynb, kb);
    }
    step_vjp(K, x, h, ynb, kb, dense, xb, cb);
}

// The terminal step alone (seed Or on r(s*), and the jet over [0, s*] when
// jet): adds the adjoints of x to xb and of (a, p_t, p_phi) to cb; returns
// the jet's sample over [0, s*].
__host__ __device__ inline double terminal_vjp(const Frame& F, const RayK& K, const double x[4], double h,
                                               double se, double Or, bool jet, double wk, double xb[4],
                                               double cb[3]) {
    double J = 0.0;
    double tmp[4] = {0.0, 0.0, 0.0, 0.0};
    step_reverse(F, K, x, h, jet, true, se, Or, wk, tmp, cb, &J);
    for (int c = 0; c < 4; ++c) xb[c] += tmp[c];
    return J;
}

// ---- one ray, forward ----------------------------------------------------

// Pixel and sample of ray index `ray` (ray = (y W + x) S^2 + k).
__host__ __device__ inline void ray_pixel(const RaygenArgs& G, int64_t ray, int& px, int& py, int& k) {
    const int S2 = G.S * G.S;
    const int64_t pix = ray / S2;
    k = static_cast<int>(ray - pix * S2);
    px = static_cast<int>(pix % G.W);
    py = static_cast<int>(pix / G.W);
}

// The camera photon of a ray and the integrator's start (src/render.cpp:
// 76-86); checkpoint 0 is written to ck.
template <class Ck>
__host__ __device__ inline void fwd_init(const Frame& F, const FrameCam& cam, int64_t ray, Ray& s, Ck& ck) {
    const RaygenArgs& G = cam.G;
    int px, py, k;
    ray_pixel(G, ray, px, py, k);
    double d[3];
    camera_dir(G, px, py, k, d);
    double pt, pr, pth, pph;
    camera_photon(G, d, pt, pr, pth, pph);
    s.y[0] = 0.0; s.y[1] = G.cam_r; s.y[2] = G.cam_th; s.y[3] = G.cam_ph; s.y[4] = pr; s.y[5] = pth;
    s.pt = pt;
    s.pph = pph;
    Rhs f;
    kerr_rhs(ray_k(F.a, F.a2, pt, pph), s.y[1], s.y[2], s.y[4], s.y[5], f);
    for (int c = 0; c < 6; ++c) s.k1[c] = f.d[c];
    s.c_prev = f.cs;
    s.h = cam.h0;
    s.lpw = kBeta * kLog1em4;   // err_prev = 1e-4 (geodesic.h:197)
    s.jet = 0.0;
    s.jet_wait = 0;
    s.nsteps = 0;
    s.tries = 0;
    if (F.K > 1)
        if (double* r = ck.put(0)) write_ck(r, s);
}

// Dense-output coefficients of the components (r, theta, p_r, p_theta).
__host__ __device__ inline void dense4(const Ray& s, const Att& A, double cf[4][5]) {
    dense_coeffs(s, A, 1, cf[0]);
    dense_coeffs(s, A, 2, cf[1]);
    dense_coeffs(s, A, 4, cf[2]);
    dense_coeffs(s, A, 5, cf[3]);
}

// One attempt of the ray in s.  Returns its class once it ends (else -1),
// with the end state in yo and, for a disc hit, the crossing fraction; the
// jet integral accumulates in s.jet (render.cpp:107-122, 150-152: sampled on
// every accepted step near the jet, over the part of the step before the
// event that ends the ray).  Checkpoints go to ck every K accepted steps.
// With pend (a 17-double slot), a step's jet sample is not taken here: its
// dense output (r, theta, p_r), s_end and h go to pend and s.jet_wait is set;
// the caller adds j0 h s_end sum_j w_j jet_value_node(j) / kConeSub to s.jet.
template <class Ck>
__host__ __device__ inline int fwd_step(const Frame& F, Ray& s, Ck& ck, double yo[6], double* s_hit,
                                        double* pend = nullptr) {
    clamp_h(F, s);
    const RayK K = ray_k(F.a, F.a2, s.pt, s.pph);
    Att A;
    attempt(K, s, F.rtol, F.atol, A);
    double lg;
    const double fac = control_fac(s, A.u, &lg);
    if (accepted(A)) {
        ++s.nsteps;
        s.tries = 0;
        const bool near = near_jet(F, s, A);
        if (F.K == 1)
            if (double* r = ck.put(s.nsteps - 1)) write_step(r, s, near);
        int term = -1;
        double s_jet = 1.0;
        const bool horizon = A.yn[1] <= F.r_stop;                    // render.cpp:172-177
        const bool disc_cross = !horizon && s.c_prev * A.g7.cs < 0.0;  // render.cpp:180-199
        if (horizon) {
            for (int c = 0; c < 6; ++c) yo[c] = A.yn[c];
            term = BHX_HORIZON;
        }
        if (disc_cross) {
            double cr[5], ct[5];
            dense_coeffs(s, A, 1, cr);
            dense_coeffs(s, A, 2, ct);
            const double se = root_theta(ct[0], ct[1], ct[2], ct[3], ct[4], s.c_prev, s.y[2], A.yn[2]);
            const double rc = dense_comp(cr[0], cr[1], cr[2], cr[3], cr[4], se);
            if (rc >= F.r_in && rc <= F.r_out) {
                for (int c = 0; c < 6; ++c) {
                    double cc[5];
                    dense_coeffs(s, A, c, cc);
                    yo[c] = dense_comp(cc[0], cc[1], cc[2], cc[3], cc[4], se);
                }
                *s_hit = se;
                s_jet = se;
                term = BHX_DISC;
            }
        }
        if (near) {
            double cf[4][5];
            dense4(s, A, cf);
            if (pend) {
                for (int c = 0; c < 3; ++c)
                    for (int m = 0; m < 5; ++m) pend[5 * c + m] = cf[c][m];
                pend[15] = s_jet;
                pend[16] = s.h;
                s.jet_wait = 1;
            } else {
                s.jet += jet_step(F, cf, s.h, s_jet, s.pt, s.pph);
            }
        }
        if (term < 0 && A.yn[1] > F.r_escape) {                       // render.cpp:247-251
            for (int c = 0; c < 6; ++c) yo[c] = A.yn[c];
            term = BHX_SKY;
        }
        if (term < 0) {
            advance(s, A, fac, lg);
            if (s.nsteps >= F.max_steps) {
                for (int c = 0; c < 6; ++c) yo[c] = s.y[c];
                term = BHX_EXHAUSTED;
            } else if (F.K > 1 && s.nsteps % F.K == 0) {
                if (double* r = ck.put(s.nsteps / F.K)) write_ck(r, s);
            }
        }
        return term;
    }
    s.h *= dmin(fac, 1.0);                                            // geodesic.h:192
    if (++s.tries > 40 || fabs(s.h) < 1e-14) {                        // render.cpp:133-136
        for (int c = 0; c < 6; ++c) yo[c] = s.y[c];
        return BHX_HORIZON;
    }
    return -1;
}

// ---- one ray, backward ---------------------------------------------------

constexpr int kSeg = 6;   // per stored step: r, theta, p_r, p_theta, h, jet-cull flag
// A lane's segment buffer holds K + 1 steps: a segment and the terminal step.

// Replay from the checkpoint at ck (the state after n_from accepted steps)
// until n_to accepted steps: each accepted step's (r, theta, p_r, p_theta, h,
// near-jet flag) goes to seg[(i * kSeg + c) * stride].  Same arithmetic as
// fwd_step, so the same decisions and the same states.
__host__ __device__ inline void replay(const Frame& F, Ray& s, const double* ck, int n_from, int n_to, double* seg,
                                       int64_t stride) {
    for (int c = 0; c < 6; ++c) s.y[c] = ck[c];
    s.h = ck[6];
    s.lpw = ck[7];
    s.nsteps = n_from;
    s.tries = 0;
    const RayK K = ray_k(F.a, F.a2, s.pt, s.pph);
    Rhs f;
    kerr_rhs(K, s.y[1], s.y[2], s.y[4], s.y[5], f);
    for (int c = 0; c < 6; ++c) s.k1[c] = f.d[c];
    s.c_prev = f.cs;
    while (s.nsteps < n_to) {
        clamp_h(F, s);
        Att A;
        attempt(K, s, F.rtol, F.atol, A);
        double lg;
        const double fac = control_fac(s, A.u, &lg);
        if (accepted(A)) {
            double* o = seg + static_cast<int64_t>(s.nsteps - n_from) * kSeg * stride;
            o[0] = s.y[1]; o[stride] = s.y[2]; o[2 * stride] = s.y[4]; o[3 * stride] = s.y[5];
            o[4 * stride] = s.h;
            o[5 * stride] = near_jet(F, s, A) ? 1.0 : 0.0;
            ++s.nsteps;
            s.tries = 0;
            advance(s, A, fac, lg);
        } else {
            s.h *= dmin(fac, 1.0);
            ++s.tries;
        }
    }
}

__host__ __device__ inline bool same_ck(const Ray& s, const double* ck) {
    for (int c = 0; c < 6; ++c)
        if (s.y[c] != ck[c]) return false;
    return s.h == ck[6] && s.lpw == ck[7];
}

// The camera end of a ray's adjoint: x0 = (r_cam, theta_cam, p_r, p_theta)
// and (p_t, p_phi) as functions of (spin, r_cam, theta_cam) through the
// frame's tetrad and metric; adds d(w I)/d(spin, theta_cam, eddington, r_cam)
// to grad from the adjoints xb of x0 and cb of (a, p_t, p_phi).
__host__ __device__ inline void bwd_camera(const Frame& F, const FrameCam& cam, const double d[3], double w,
                                           double wk, double I_disc, double J, const double xb[4],
                                           const double cb[3], double grad[4]) {
    const RaygenArgs& G = cam.G;
    using D3 = Dual<3>;
    const double base[10] = {G.e00, G.e03, G.e11, G.e22, G.e33, G.gl_tt, G.gl_tp, G.gl_rr, G.gl_hh, G.gl_pp};
    D3 e[10];
    for (int i = 0; i < 10; ++i) {
        e[i] = D3(base[i]);
        for (int m = 0; m < 3; ++m) e[i].d[m] = cam.dcam[i][m];
    }
    D3 pt, pr, pth, pph;
    camera_photon_t<D3>(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], d, pt, pr, pth, pph);
    double dp[3];
    for (int m = 0; m < 3; ++m)
        dp[m] = xb[2] * pr.d[m] + xb[3] * pth.d[m] + cb[1] * pt.d[m] + cb[2] * pph.d[m];
    const double I = I_disc + F.kappa * J;
    grad[0] += cb[0] + dp[0] + wk * J * F.dlnj0_da;   // spin (incl. the jet power's)
    grad[1] += xb[1] + dp[2];                         // theta_cam
    grad[2] += w * I / F.eddington;                   // eddington: disc flux and jet power ~ Mdot
    grad[3] += xb[0] + dp[1];                         // r_cam
}

// The adjoint of one ray from a record of every accepted step (K = 1): no
// replay.  The records hold what a replay would give - each step's start
// state, accepted step size and near-jet flag, written by the same forward
// arithmetic - so the result equals bwd_ray's at any other K bit for bit.
// A disc hit's terminal step is taken again from its record for the crossing
// fraction and radius.  Returns 1 when that step does not reproduce the
// forward's crossing (0 otherwise).
template <class Ck>
__host__ __device__ inline int bwd_ray_log(const Frame& F, const FrameCam& cam, int64_t ray, int n_acc, int cls,
                                           double w, const Ck& ck, double grad[4], double* lds = nullptr) {
    const RaygenArgs& G = cam.G;
    int mismatch = 0;
    int px, py, k;
    ray_pixel(G, ray, px, py, k);
    double d[3];
    camera_dir(G, px, py, k, d);
    Ray s;
    {
        double pr, pth;
        camera_photon(G, d, s.pt, pr, pth, s.pph);
    }
    const RayK K = ray_k(F.a, F.a2, s.pt, s.pph);
    const double wk = w * F.kappa;
    const bool disc = cls == BHX_DISC;
    double xb[4] = {0.0, 0.0, 0.0, 0.0};
    double cb[3] = {0.0, 0.0, 0.0};
    double I_disc = 0.0, J = 0.0, se = 0.5, Or = 0.0;
    CkPos p = ck.rpos(n_acc);
    if (disc) {
        // t and phi (not recorded) enter only the step's error norm, not its
        // r and theta: the crossing comes out as in the forward
        const double* o = ck.at(p);
        s.y[0] = 0.0; s.y[1] = o[0]; s.y[2] = o[1]; s.y[3] = 0.0; s.y[4] = o[2]; s.y[5] = o[3];
        s.h = o[4];
        Rhs f;
        kerr_rhs(K, s.y[1], s.y[2], s.y[4], s.y[5], f);
        for (int c = 0; c < 6; ++c) s.k1[c] = f.d[c];
        s.c_prev = f.cs;
        Att A;
        attempt(K, s, F.rtol, F.atol, A);
        double cr[5], ct[5];
        dense_coeffs(s, A, 1, cr);
        dense_coeffs(s, A, 2, ct);
        se = root_theta(ct[0], ct[1], ct[2], ct[3], ct[4], s.c_prev, s.y[2], A.yn[2]);
        const double r_hit = dense_comp(cr[0], cr[1], cr[2], cr[3], cr[4], se);
        if (!(s.c_prev * A.g7.cs < 0.0) || !(r_hit >= F.r_in && r_hit <= F.r_out)) ++mismatch;
        I_disc = disc_intensity(F, r_hit, s.pt, s.pph);
        if (I_disc > 0.0) {
            double dI[4];
            disc_intensity_grad(F, r_hit, s.pt, s.pph, dI);
            Or = w * dI[0];
            cb[0] += w * dI[1]; cb[1] += w * dI[2]; cb[2] += w * dI[3];
        }
    }
    // sweep back over the records, the terminal step first; record i - 1 is
    // loaded while step i is swept
    double nx[kSeg];
    {
        const double* o = ck.at(p);
        for (int c = 0; c < kSeg; ++c) nx[c] = o[c];
    }
    for (int i = n_acc - 1; i >= 0; --i) {
        const double xi[4] = {nx[0], nx[1], nx[2], nx[3]};
        const double hi = nx[4];
        const bool jet_i = nx[5] != 0.0;
        if (i > 0) {
            ck.step_back(p);
            const double* o = ck.at(p);
            for (int c = 0; c < kSeg; ++c) nx[c] = o[c];
        }
        const bool term = disc && i == n_acc - 1;
        step_reverse(F, K, xi, hi, jet_i, term, se, Or, wk, xb, cb, &J, lds);
    }
    bwd_camera(F, cam, d, w, wk, I_disc, J, xb, cb, grad);
    return mismatch;
}

// Adjoint of w * I for one ray that ended after n_acc accepted steps with
// class cls, I = sigma (g T)^4 / pi of a disc hit + kappa J of the jet:
// adds d(w I)/d(spin, theta_cam, eddington, r_cam) to grad.  Returns the
// number of replayed segments whose end state differed from the next
// checkpoint (0 when the replay reproduces the forward exactly).
template <class Ck>
__host__ __device__ inline int bwd_ray(const Frame& F, const FrameCam& cam, int64_t ray, int n_acc, int cls,
                                       double w, const Ck& ck, double* seg, int64_t stride, double grad[4],
                                       double* lds = nullptr) {
    if (F.K == 1) return bwd_ray_log(F, cam, ray, n_acc, cls, w, ck, grad, lds);
    const RaygenArgs& G = cam.G;
    int mismatch = 0;
    int px, py, k;
    ray_pixel(G, ray, px, py, k);
    double d[3];
    camera_dir(G, px, py, k, d);
    Ray s;
    {
        double pr, pth;
        camera_photon(G, d, s.pt, pr, pth, s.pph);
    }
    const RayK K = ray_k(F.a, F.a2, s.pt, s.pph);
    const double wk = w * F.kappa;       // weight of the jet integral in w I
    const bool disc = cls == BHX_DISC;
    double xb[4] = {0.0, 0.0, 0.0, 0.0};
    double cb[3] = {0.0, 0.0, 0.0};
    double I_disc = 0.0, J = 0.0, se = 0.5, Or = 0.0;
    // the last segment: replayed up to the terminal step (a disc hit: stored
    // after the segment's steps) or to the end (every other class: all steps
    // are full steps)
    const int n_reg = disc ? n_acc - 1 : n_acc;
    const int j_last = (n_reg == 0) ? 0 : (n_reg - 1) / F.K;
    const int last_from = j_last * F.K;
    replay(F, s, ck.get(j_last), last_from, n_reg, seg, stride);
    if (disc) {
        double r_hit = 0.0;
        bool jet_term = false;
        for (;;) {
            clamp_h(F, s);
            Att A;
            attempt(K, s, F.rtol, F.atol, A);
            double lg;
            const double fac = control_fac(s, A.u, &lg);
            if (accepted(A)) {
                double cr[5], ct[5];
                dense_coeffs(s, A, 1, cr);
                dense_coeffs(s, A, 2, ct);
                se = root_theta(ct[0], ct[1], ct[2], ct[3], ct[4], s.c_prev, s.y[2], A.yn[2]);
                r_hit = dense_comp(cr[0], cr[1], cr[2], cr[3], cr[4], se);
                if (!(s.c_prev * A.g7.cs < 0.0) || !(r_hit >= F.r_in && r_hit <= F.r_out)) ++mismatch;
                jet_term = near_jet(F, s, A);
                break;
            }
            s.h *= dmin(fac, 1.0);
            if (++s.tries > 40) { ++mismatch; return mismatch; }
        }
        double* o = seg + static_cast<int64_t>(n_reg - last_from) * kSeg * stride;
        o[0] = s.y[1]; o[stride] = s.y[2]; o[2 * stride] = s.y[4]; o[3 * stride] = s.y[5];
        o[4 * stride] = s.h;
        o[5 * stride] = jet_term ? 1.0 : 0.0;
        I_disc = disc_intensity(F, r_hit, s.pt, s.pph);
        if (I_disc > 0.0) {
            double dI[4];
            disc_intensity_grad(F, r_hit, s.pt, s.pph, dI);
            Or = w * dI[0];
            cb[0] += w * dI[1]; cb[1] += w * dI[2]; cb[2] += w * dI[3];
        }
    }
    // sweep back through the stored segments (the terminal step first)
    const int n_last = disc ? n_reg + 1 : n_reg;
    for (int j = j_last; j >= 0; --j) {
        const int n_from = j * F.K;
        const int n_to = (j == j_last) ? n_last : n_from + F.K;
        if (j != j_last) {
            replay(F, s, ck.get(j), n_from, n_to, seg, stride);
            if (!same_ck(s, ck.get(j + 1))) ++mismatch;
        }
        // the stored step i - 1 is loaded while step i is swept (one wave per
        // SIMD hides no load latency by itself)
        double nx[kSeg];
        {
            const double* o = seg + static_cast<int64_t>(n_to - n_from - 1) * kSeg * stride;
            for (int c = 0; c < kSeg; ++c) nx[c] = o[c * stride];
        }
        for (int i = n_to - n_from - 1; i >= 0; --i) {
            const double xi[4] = {nx[0], nx[1], nx[2], nx[3]};
            const double hi = nx[4];
            const bool jet_i = nx[5] != 0.0;
            if (i > 0) {
                const double* o = seg + static_cast<int64_t>(i - 1) * kSeg * stride;
                for (int c = 0; c < kSeg; ++c) nx[c] = o[c * stride];
            }
            const bool term = disc && j == j_last && i == n_to - n_from - 1;
            step_reverse(F, K, xi, hi, jet_i, term, se, Or, wk, xb, cb, &J, lds);
        }
    }
    bwd_camera(F, cam, d, w, wk, I_disc, J, xb, cb, grad);
    return mismatch;
}

}  // namespace adj
}  // namespace bhx
