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
constexpr int kCk = 8;   // checkpoint record: y[6], h, controller memory

// Per-frame constants of a differentiable render.
struct Frame {
    // integrator (src/render.cpp:71-86)
    double a, a2, r_h, r_stop, rtol, atol, r_escape, h0, pw_init, r_in, r_out;
    int32_t max_steps, K, cap, sg;
    // Novikov-Thorne disc (src/disc.h:46-121) and the spin derivatives of its
    // spin-only constants
    double disc_a, flux_scale, dflux_da, eddington;
    double pt_x0, pt_c[3], pt_x[3], d_x0, d_c[3], d_x[3];
    // d(e00 e03 e11 e22 e33 gl_tt gl_tp gl_rr gl_hh gl_pp)/d(a, r_cam, theta_cam)
    double dcam[10][3];
};

// Integrator state of one ray.
struct Ray {
    double y[6], k1[6], pt, pph, h, pw, c_prev;
    int nsteps, tries;
};

__host__ __device__ inline void write_ck(double* ck, const Ray& s) {
    for (int c = 0; c < 6; ++c) ck[c] = s.y[c];
    ck[6] = s.h;
    ck[7] = s.pw;
}

// One DP5 attempt from s (the tracer's arithmetic, src/geodesic.h:130-169).
struct Att {
    double yn[6], c5[6], err;
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
        const double sc = atol + rtol * dmax(fabs(y[c]), fabs(o.yn[c]));
        const double qe = e * rcp(sc);
        err2 += qe * qe;
        o.c5[c] = h * (D1 * k1[c] + D3 * k3[c] + D4 * k4[c] + D5 * k5[c] + D6 * k6[c] + D7 * o.g7.d[c]);
    }
    double err = sqrt(err2 * 0.125);
    if (!finite_d(err)) err = 1e10;
    o.err = err;
}

// hmax clamp before a step (src/render.cpp:130-131), once per step.
__host__ __device__ inline void clamp_h(const Frame& F, Ray& s) {
    const double hmax = 0.05 * dmax(1.0, s.y[1] - F.r_h) + 0.01;
    if (s.tries == 0 && s.h > hmax) s.h = hmax;
}

// PI controller factor (src/geodesic.h:170-173) and the log of the error.
__host__ __device__ inline double control_fac(const Ray& s, double err, double* lg_out) {
    const double lg = log_pos(dmax(err, 1e-16));
    double fac = 0.9 * exp_near(-kAlpha * lg) * s.pw;
    *lg_out = lg;
    return dclamp(fac, 0.2, 5.0);
}

// Accepted step that does not end the ray: advance the state.
__host__ __device__ inline void advance(Ray& s, const Att& A, double fac, double lg) {
    for (int c = 0; c < 6; ++c) { s.y[c] = A.yn[c]; s.k1[c] = A.g7.d[c]; }
    s.c_prev = A.g7.cs;
    s.pw = exp_near(kBeta * dmax(lg, kLog1em4));
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
    const double iS = 1.0 / S, iD = 1.0 / D, is2 = 1.0 / s2;
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

// Reverse of the terminal step: r_hit = r(s*) on the dense output of the step
// from x with size h, where theta(s*) = pi/2.  Seeds the adjoint jr of r_hit;
// adds the adjoint of x to xb and of (a, p_t, p_phi) to cb.
__host__ __device__ inline void terminal_vjp(const RayK& K, const double x[4], double h, double s, double jr,
                                             double xb[4], double cb[3]) {
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
    // dense coefficients of r (0) and theta (1), their s-derivatives at s*
    double dv[2];
    double cf[2][5];
    for (int c = 0; c < 2; ++c) {
        cf[c][0] = x[c];
        cf[c][1] = yn[c] - x[c];
        cf[c][2] = h * k[0][c] - cf[c][1];
        cf[c][3] = cf[c][1] - h * k[6][c] - cf[c][2];
        double acc = 0.0;
        for (int j = 0; j < 7; ++j) acc += Dd[j] * k[j][c];
        cf[c][4] = h * acc;
        double v;
        dense_vd(cf[c][0], cf[c][1], cf[c][2], cf[c][3], cf[c][4], s, v, dv[c]);
    }
    // implicit function theorem: d r_hit = d r(s*) - r'(s*) / theta'(s*) d theta(s*)
    const double yb[2] = {jr, -jr * dv[0] / dv[1]};
    const double s1 = 1.0 - s;
    const double wts[5] = {1.0, s, s * s1, s * s * s1, s * s * s1 * s1};
    double kb[7][4] = {};
    double ynb[4] = {0.0, 0.0, 0.0, 0.0};
    for (int c = 0; c < 2; ++c) {
        double cbr[5];
        for (int m = 0; m < 5; ++m) cbr[m] = wts[m] * yb[c];
        for (int j = 0; j < 7; ++j) kb[j][c] += h * Dd[j] * cbr[4];    // c5 = h sum d_j k_j
        cbr[1] += cbr[3]; kb[6][c] += -h * cbr[3]; cbr[2] += -cbr[3];  // c4 = c2 - h k7 - c3
        kb[0][c] += h * cbr[2]; cbr[1] += -cbr[2];                     // c3 = h k1 - c2
        ynb[c] += cbr[1]; xb[c] += -cbr[1];                            // c2 = yn - x
        xb[c] += cbr[0];                                               // c1 = x
    }
    step_vjp(K, x, h, ynb, kb, true, xb, cb);
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
__host__ __device__ inline void fwd_init(const Frame& F, const RaygenArgs& G, int64_t ray, Ray& s, double* ck) {
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
    s.h = F.h0;
    s.pw = F.pw_init;
    s.nsteps = 0;
    s.tries = 0;
    write_ck(ck, s);
}

// One attempt of the ray in s.  Returns its class once it ends (else -1),
// with the end state in yo and, for a disc hit, the crossing fraction.
// Checkpoints go to ck (this ray's slots) every K accepted steps.
__host__ __device__ inline int fwd_step(const Frame& F, Ray& s, double* ck, double yo[6], double* s_hit) {
    clamp_h(F, s);
    const RayK K = ray_k(F.a, F.a2, s.pt, s.pph);
    Att A;
    attempt(K, s, F.rtol, F.atol, A);
    double lg;
    const double fac = control_fac(s, A.err, &lg);
    if (A.err <= 1.0) {
        ++s.nsteps;
        s.tries = 0;
        int term = -1;
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
                term = BHX_DISC;
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
            } else if (s.nsteps % F.K == 0) {
                write_ck(ck + (s.nsteps / F.K) * kCk, s);
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

// Replay from the checkpoint at ck (the state after n_from accepted steps)
// until n_to accepted steps: each accepted step's (r, theta, p_r, p_theta, h)
// goes to seg[(i * 5 + c) * stride].  Same arithmetic as fwd_step, so the
// same decisions and the same states.
__host__ __device__ inline void replay(const Frame& F, Ray& s, const double* ck, int n_from, int n_to, double* seg,
                                       int64_t stride) {
    for (int c = 0; c < 6; ++c) s.y[c] = ck[c];
    s.h = ck[6];
    s.pw = ck[7];
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
        const double fac = control_fac(s, A.err, &lg);
        if (A.err <= 1.0) {
            double* o = seg + static_cast<int64_t>(s.nsteps - n_from) * 5 * stride;
            o[0] = s.y[1]; o[stride] = s.y[2]; o[2 * stride] = s.y[4]; o[3 * stride] = s.y[5];
            o[4 * stride] = s.h;
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
    return s.h == ck[6] && s.pw == ck[7];
}

// Adjoint of w * I for one disc-hit ray that ended after n_acc accepted
// steps: adds d(w I)/d(spin, theta_cam, eddington, r_cam) to grad.  Returns
// the number of replayed segments whose end state differed from the next
// checkpoint (0 when the replay reproduces the forward exactly).
__host__ __device__ inline int bwd_ray(const Frame& F, const RaygenArgs& G, int64_t ray, int n_acc, double w,
                                       const double* ck, double* seg, int64_t stride, double grad[4]) {
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
    const int n_term = n_acc - 1;           // accepted steps before the terminal one
    const int j_last = n_term / F.K;
    // the terminal segment up to the terminal step, then the terminal step itself
    replay(F, s, ck + j_last * kCk, j_last * F.K, n_term, seg, stride);
    double xt[4] = {s.y[1], s.y[2], s.y[4], s.y[5]};
    double se = 0.5, r_hit = 0.0;
    for (;;) {
        clamp_h(F, s);
        Att A;
        attempt(K, s, F.rtol, F.atol, A);
        double lg;
        const double fac = control_fac(s, A.err, &lg);
        if (A.err <= 1.0) {
            double cr[5], ct[5];
            dense_coeffs(s, A, 1, cr);
            dense_coeffs(s, A, 2, ct);
            se = root_theta(ct[0], ct[1], ct[2], ct[3], ct[4], s.c_prev, s.y[2], A.yn[2]);
            r_hit = dense_comp(cr[0], cr[1], cr[2], cr[3], cr[4], se);
            if (!(s.c_prev * A.g7.cs < 0.0) || !(r_hit >= F.r_in && r_hit <= F.r_out)) ++mismatch;
            break;
        }
        s.h *= dmin(fac, 1.0);
        if (++s.tries > 40) { ++mismatch; return mismatch; }
    }
    const double h_term = s.h;
    // intensity and its explicit partials at the hit
    const double I = disc_intensity(F, r_hit, s.pt, s.pph);
    if (!(I > 0.0)) return mismatch;
    double dI[4];
    disc_intensity_grad(F, r_hit, s.pt, s.pph, dI);
    double xb[4] = {0.0, 0.0, 0.0, 0.0};
    double cb[3] = {w * dI[1], w * dI[2], w * dI[3]};
    terminal_vjp(K, xt, h_term, se, w * dI[0], xb, cb);
    // sweep back through the stored segments
    for (int j = j_last; j >= 0; --j) {
        const int n_from = j * F.K;
        const int n_to = (j == j_last) ? n_term : n_from + F.K;
        if (j != j_last) {
            replay(F, s, ck + j * kCk, n_from, n_to, seg, stride);
            if (!same_ck(s, ck + (j + 1) * kCk)) ++mismatch;
        }
        for (int i = n_to - n_from - 1; i >= 0; --i) {
            const double* o = seg + static_cast<int64_t>(i) * 5 * stride;
            const double xi[4] = {o[0], o[stride], o[2 * stride], o[3 * stride]};
            double ynb[4] = {xb[0], xb[1], xb[2], xb[3]};
            double kb[7][4] = {};
            xb[0] = xb[1] = xb[2] = xb[3] = 0.0;
            step_vjp(K, xi, o[4 * stride], ynb, kb, false, xb, cb);
        }
    }
    // camera: x0 = (r_cam, theta_cam, p_r, p_theta) and (p_t, p_phi) as
    // functions of (spin, r_cam, theta_cam) through the frame's tetrad and metric
    using D3 = Dual<3>;
    const double base[10] = {G.e00, G.e03, G.e11, G.e22, G.e33, G.gl_tt, G.gl_tp, G.gl_rr, G.gl_hh, G.gl_pp};
    D3 e[10];
    for (int i = 0; i < 10; ++i) {
        e[i] = D3(base[i]);
        for (int m = 0; m < 3; ++m) e[i].d[m] = F.dcam[i][m];
    }
    D3 pt, pr, pth, pph;
    camera_photon_t<D3>(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7], e[8], e[9], d, pt, pr, pth, pph);
    double dp[3];
    for (int m = 0; m < 3; ++m)
        dp[m] = xb[2] * pr.d[m] + xb[3] * pth.d[m] + cb[1] * pt.d[m] + cb[2] * pph.d[m];
    grad[0] += cb[0] + dp[0];           // spin
    grad[1] += xb[1] + dp[2];           // theta_cam
    grad[2] += w * I / F.eddington;     // eddington (F is linear in the accretion rate)
    grad[3] += xb[0] + dp[1];           // r_cam
    return mismatch;
}

}  // namespace adj
}  // namespace bhx
