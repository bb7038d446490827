// oracle.cpp - CPU oracle of the per-pixel Kerr tracer.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
// of bench.py.  Nothing in blackhole_amd/ includes, links or loads this file.
//
// A restatement of MattPar-Cell/Blackhole's tracer from its algorithm.  Every
// arithmetic expression keeps the reference's association and operand order
// (cited per function) so that, compiled with -O3 -ffp-contract=off on the
// same glibc, results are bitwise equal to the reference built from source.
// The code is organised differently (plain arrays, free functions, one
// struct of precomputed scene constants) and shares no text with it.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---- physical constants (src/constants.h:23-45) --------------------------
constexpr double kG = 6.67430e-11;
constexpr double kC = 2.99792458e8;
constexpr double kH = 6.62607015e-34;
constexpr double kKB = 1.380649e-23;
constexpr double kSigmaSB = 5.670374419e-8;
constexpr double kSigmaT = 6.6524587321e-29;
constexpr double kMp = 1.67262192369e-27;
constexpr double kMsun = 1.98847e30;
constexpr double kRsun = 6.957e8;

double rg_metres(double M) { return kG * M / (kC * kC); }                    // constants.h:52
double eddington_lum(double M) { return 4.0 * kPi * kG * M * kMp * kC / kSigmaT; }  // :56-58
double eddington_mdot(double M, double eta) { return eddington_lum(M) / (eta * kC * kC); }  // :62-64

// ---- Kerr geometry (src/kerr.h:37-176) ------------------------------------
// The geometry, the Hamiltonian flow, the DP5 attempt and the per-step
// invariants are templates over the scalar type: double everywhere in the
// oracle, and a FLOP-counting type for oracle_flop_counts (the algorithmic
// operation counts of the roofline).  T = double compiles to the same
// arithmetic as plain functions.
template <class T>
struct PtT {                   // shared sub-expressions at (r, theta)
    T a, a2, r, sn, cs, sn2, cs2, S, D, A;
};
using Pt = PtT<double>;

// std::max / std::min / std::clamp with the same comparisons, for any T
template <class T, class U> T tmax(const T& x, const U& y) { return (x < y) ? T(y) : x; }
template <class T, class U> T tmin(const T& x, const U& y) { return (y < x) ? T(y) : x; }
template <class T> T tclamp(const T& v, double lo, double hi) { return (v < lo) ? T(lo) : ((hi < v) ? T(hi) : v); }

template <class T>
PtT<T> point_t(T a, T r, T th) {  // kerr.h:44-59 (sin clamp kerr.h:34,51)
    using std::cos; using std::fabs; using std::sin;
    PtT<T> p;
    p.a = a;
    p.a2 = a * a;
    p.r = r;
    p.sn = sin(th);
    if (fabs(p.sn) < 1e-9) p.sn = (p.sn < 0.0) ? T(-1e-9) : T(1e-9);
    p.cs = cos(th);
    p.sn2 = p.sn * p.sn;
    p.cs2 = p.cs * p.cs;
    p.S = r * r + p.a2 * p.cs2;
    p.D = r * r - 2.0 * r + p.a2;
    p.A = (r * r + p.a2) * (r * r + p.a2) - p.a2 * p.D * p.sn2;
    return p;
}
inline Pt point(double a, double r, double th) { return point_t<double>(a, r, th); }

// Index order of the five non-zero components: tt, tp, rr, hh (theta-theta), pp.
enum { TT = 0, TP = 1, RR = 2, HH = 3, PP = 4 };

template <class T>
void cov_metric(const PtT<T>& p, T g[5]) {  // kerr.h:69-77
    g[TT] = -(1.0 - 2.0 * p.r / p.S);
    g[TP] = -2.0 * p.a * p.r * p.sn2 / p.S;
    g[RR] = p.S / p.D;
    g[HH] = p.S;
    g[PP] = p.A * p.sn2 / p.S;
}

template <class T>
void inv_metric(const PtT<T>& p, T g[5]) {  // kerr.h:107-116
    const T SD = p.S * p.D;
    g[TT] = -p.A / SD;
    g[TP] = -2.0 * p.a * p.r / SD;
    g[RR] = p.D / p.S;
    g[HH] = 1.0 / p.S;
    g[PP] = (p.D - p.a2 * p.sn2) / (SD * p.sn2);
}

// Inverse metric and its r, theta partials, closed form (kerr.h:129-176).
template <class T>
void inv_metric_partials(const PtT<T>& p, T g[5], T gr[5], T gh[5]) {
    inv_metric(p, g);
    const T r = p.r, a2 = p.a2, a = p.a, sn = p.sn, cs = p.cs, sn2 = p.sn2;
    const T S = p.S, D = p.D, A = p.A;
    const T Sr = 2.0 * r, Sh = -2.0 * a2 * sn * cs;
    const T Dr = 2.0 * r - 2.0, Dh = 0.0;
    const T Ar = 4.0 * r * (r * r + a2) - a2 * Dr * sn2;
    const T Ah = -2.0 * a2 * D * sn * cs;
    const T Q = S * D;
    const T Qr = Sr * D + S * Dr;
    const T Qh = Sh * D + S * Dh;
    const T Q2 = Q * Q;
    gr[TT] = -(Ar * Q - A * Qr) / Q2;
    gh[TT] = -(Ah * Q - A * Qh) / Q2;
    const T N = -2.0 * a * r, Nr = -2.0 * a, Nh = 0.0;
    gr[TP] = (Nr * Q - N * Qr) / Q2;
    gh[TP] = (Nh * Q - N * Qh) / Q2;
    gr[RR] = (Dr * S - D * Sr) / (S * S);
    gh[RR] = (Dh * S - D * Sh) / (S * S);
    gr[HH] = -Sr / (S * S);
    gh[HH] = -Sh / (S * S);
    const T P = D - a2 * sn2, Pr = Dr, Ph = -2.0 * a2 * sn * cs;
    const T Rq = Q * sn2, Rr = Qr * sn2, Rh = Qh * sn2 + Q * 2.0 * sn * cs;
    gr[PP] = (Pr * Rq - P * Rr) / (Rq * Rq);
    gh[PP] = (Ph * Rq - P * Rh) / (Rq * Rq);
}

double horizon_plus(double a) { return 1.0 + std::sqrt(std::max(0.0, 1.0 - a * a)); }   // kerr.h:187-190
double horizon_minus(double a) { return 1.0 - std::sqrt(std::max(0.0, 1.0 - a * a)); }  // kerr.h:191-194

double isco(double a, int sense) {  // kerr.h:223-229 (Bardeen, Press & Teukolsky 1972)
    const double a2 = a * a;
    const double Z1 = 1.0 + std::cbrt(1.0 - a2) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    const double Z2 = std::sqrt(3.0 * a2 + Z1 * Z1);
    const double s = std::sqrt(std::max(0.0, (3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2)));
    return 3.0 + Z2 - static_cast<double>(sense) * s;
}

double circ_energy(double a, double r, int sg) {  // kerr.h:232-237
    const double sr = std::sqrt(r);
    const double num = r * r - 2.0 * r + sg * a * sr;
    const double den = r * std::sqrt(r * r - 3.0 * r + 2.0 * sg * a * sr);
    return num / den;
}

double circ_angmom(double a, double r, int sg) {  // kerr.h:240-245
    const double sr = std::sqrt(r);
    const double num = sg * sr * (r * r - 2.0 * sg * a * sr + a * a);
    const double den = r * std::sqrt(r * r - 3.0 * r + 2.0 * sg * a * sr);
    return num / den;
}

void kepler_u(double a, double r, int sg, double u[4]) {  // kerr.h:260-266
    const double r32 = r * std::sqrt(r);
    const double Om = sg / (r32 + sg * a);
    const double disc = r * r * r - 3.0 * r * r + 2.0 * sg * a * r32;
    const double ut = (r32 + sg * a) / std::sqrt(std::max(1e-300, disc));
    u[0] = ut; u[1] = 0.0; u[2] = 0.0; u[3] = Om * ut;
}

// ZAMO tetrad e[a][mu] (kerr.h:270-293): zero except the five entries set.
void zamo(const Pt& p, double e[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) e[i][j] = 0.0;
    const double alpha = std::sqrt(p.S * p.D / p.A);
    const double omega = 2.0 * p.a * p.r / p.A;
    e[0][0] = 1.0 / alpha;
    e[0][3] = omega / alpha;
    e[1][1] = std::sqrt(p.D / p.S);
    e[2][2] = 1.0 / std::sqrt(p.S);
    e[3][3] = std::sqrt(p.S / p.A) / p.sn;
}

template <class T>
T carter_Q(T a, T th, T pth, T E, T Lz) {  // kerr.h:342-348, m2 = 0
    using std::cos; using std::fabs; using std::sin;
    T sn = sin(th);
    if (fabs(sn) < 1e-9) sn = T(1e-9);
    const T cs2 = cos(th) * cos(th);
    return pth * pth + cs2 * (a * a * (0.0 - E * E) + Lz * Lz / (sn * sn));
}

// ---- Hamiltonian flow (src/geodesic.h:50-71) ------------------------------
// y = (t, r, th, ph, p_t, p_r, p_th, p_ph)
enum { YT = 0, YR = 1, YH = 2, YP = 3, YPT = 4, YPR = 5, YPH = 6, YPP = 7 };

template <class T>
void rhs_t(T a, const T y[8], T d[8]) {
    const PtT<T> p = point_t<T>(a, y[YR], y[YH]);
    T g[5], gr[5], gh[5];
    inv_metric_partials(p, g, gr, gh);
    const T pt = y[YPT], pr = y[YPR], ph = y[YPH], pp = y[YPP];
    d[YT] = g[TT] * pt + g[TP] * pp;
    d[YR] = g[RR] * pr;
    d[YH] = g[HH] * ph;
    d[YP] = g[TP] * pt + g[PP] * pp;
    d[YPT] = T(0.0);
    d[YPP] = T(0.0);
    d[YPR] = -0.5 * (gr[TT] * pt * pt + 2.0 * gr[TP] * pt * pp + gr[RR] * pr * pr +
                     gr[HH] * ph * ph + gr[PP] * pp * pp);
    d[YPH] = -0.5 * (gh[TT] * pt * pt + 2.0 * gh[TP] * pt * pp + gh[RR] * pr * pr +
                     gh[HH] * ph * ph + gh[PP] * pp * pp);
}

inline void rhs(double a, const double y[8], double d[8]) { rhs_t<double>(a, y, d); }

// ---- Dormand-Prince 5(4) coefficients (Hairer, Norsett & Wanner II.4) ----
// Same literal fractions as geodesic.h:77-106 (they are the published tableau).
namespace dp5 {
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                 a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
constexpr double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0,
                 b5 = -2187.0 / 6784.0, b6 = 11.0 / 84.0;
constexpr double e1 = b1 - 5179.0 / 57600.0, e3 = b3 - 7571.0 / 16695.0,
                 e4 = b4 - 393.0 / 640.0, e5 = b5 - (-92097.0 / 339200.0),
                 e6 = b6 - 187.0 / 2100.0, e7 = 0.0 - 1.0 / 40.0;
constexpr double d1 = -12715105075.0 / 11282082432.0, d3 = 87487479700.0 / 32700410799.0,
                 d4 = -10690763975.0 / 1880347072.0, d5 = 701980252875.0 / 199316789632.0,
                 d6 = -1453857185.0 / 822651844.0, d7 = 69997945.0 / 29380423.0;
}  // namespace dp5

// Hairer's dense-output polynomial of one accepted step (geodesic.h:109-121).
template <class T>
struct SegT {
    T c1[8], c2[8], c3[8], c4[8], c5[8];
    T h;
    void at(double s, double out[8]) const {
        const double s1 = 1.0 - s;
        for (int i = 0; i < 8; ++i)
            out[i] = c1[i] + s * (c2[i] + s1 * (c3[i] + s * (c4[i] + s1 * c5[i])));
    }
    double comp(double s, int i) const {
        const double s1 = 1.0 - s;
        return c1[i] + s * (c2[i] + s1 * (c3[i] + s * (c4[i] + s1 * c5[i])));
    }
};
using Seg = SegT<double>;

// One integrator instance: FSAL first stage + PI-controller memory
// (geodesic.h:123-204).
inline bool finite_t(double x) { return std::isfinite(x); }

template <class T>
struct DP5T {
    T a, rtol, atol;
    T k1[8];
    bool k1_ok = false;
    T err_prev = T(1e-4);
    long long rhs_calls = 0;

    void reset() { k1_ok = false; err_prev = T(1e-4); }

    bool attempt(const T y[8], T& h, T ynew[8], SegT<T>& seg) {
        using namespace dp5;
        using std::fabs; using std::pow; using std::sqrt;
        if (!k1_ok) { rhs_t<T>(a, y, k1); k1_ok = true; ++rhs_calls; }
        T t[8], k2[8], k3[8], k4[8], k5[8], k6[8], k7[8];
        for (int i = 0; i < 8; ++i) t[i] = y[i] + h * a21 * k1[i];
        rhs_t<T>(a, t, k2);
        for (int i = 0; i < 8; ++i) t[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
        rhs_t<T>(a, t, k3);
        for (int i = 0; i < 8; ++i) t[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
        rhs_t<T>(a, t, k4);
        for (int i = 0; i < 8; ++i)
            t[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
        rhs_t<T>(a, t, k5);
        for (int i = 0; i < 8; ++i)
            t[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
        rhs_t<T>(a, t, k6);
        for (int i = 0; i < 8; ++i)
            ynew[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
        rhs_t<T>(a, ynew, k7);
        rhs_calls += 6;

        T acc = T(0.0);
        for (int i = 0; i < 8; ++i) {
            const T e = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] +
                             e7 * k7[i]);
            const T sc = atol + rtol * tmax(fabs(y[i]), fabs(ynew[i]));
            const T q = e / sc;
            acc += q * q;
        }
        T err = sqrt(acc / 8.0);
        if (!finite_t(err)) err = T(1e10);

        const double alpha = 0.7 / 5.0, beta = 0.4 / 5.0, safety = 0.9;
        T fac = safety * pow(tmax(err, 1e-16), -alpha) *
                pow(tmax(err_prev, 1e-4), beta);
        fac = tclamp(fac, 0.2, 5.0);

        if (err <= 1.0) {
            seg.h = h;
            for (int i = 0; i < 8; ++i) {
                seg.c1[i] = y[i];
                seg.c2[i] = ynew[i] - y[i];
                seg.c3[i] = h * k1[i] - seg.c2[i];
                seg.c4[i] = seg.c2[i] - h * k7[i] - seg.c3[i];
                seg.c5[i] = h * (d1 * k1[i] + d3 * k3[i] + d4 * k4[i] + d5 * k5[i] + d6 * k6[i] +
                                 d7 * k7[i]);
            }
            for (int i = 0; i < 8; ++i) k1[i] = k7[i];
            err_prev = tmax(err, 1e-4);
            h *= fac;
            return true;
        }
        h *= tmin(fac, 1.0);
        return false;
    }
};
using DP5 = DP5T<double>;

// Null momentum from a local direction in the ZAMO frame, past-directed
// (geodesic.h:215-231 with past_directed = true).
void photon(double a, double r, double th, double phi, const double e[4][4], const double dir[3],
            double y[8]) {
    const Pt p = point(a, r, th);
    const double s = -1.0;
    double pu[4];
    for (int mu = 0; mu < 4; ++mu)
        pu[mu] = e[0][mu] + s * (dir[0] * e[1][mu] + dir[1] * e[2][mu] + dir[2] * e[3][mu]);
    for (int mu = 0; mu < 4; ++mu) pu[mu] = -pu[mu];
    double g[5];
    cov_metric(p, g);
    y[YT] = 0.0; y[YR] = r; y[YH] = th; y[YP] = phi;
    y[YPT] = g[TT] * pu[0] + g[TP] * pu[3];
    y[YPR] = g[RR] * pu[1];
    y[YPH] = g[HH] * pu[2];
    y[YPP] = g[TP] * pu[0] + g[PP] * pu[3];
}

template <class T>
struct InvT { T E, Lz, Q, norm, norm_rel; };
using Inv = InvT<double>;

template <class T>
InvT<T> invariants_t(T a, const T y[8]) {  // geodesic.h:240-251, kerr.h:301-321
    using std::fabs;
    InvT<T> v;
    v.E = -y[YPT];
    v.Lz = y[YPP];
    v.Q = carter_Q<T>(a, y[YH], y[YPH], v.E, v.Lz);
    T g[5];
    inv_metric(point_t<T>(a, y[YR], y[YH]), g);
    const T p0 = y[YPT], p1 = y[YPR], p2 = y[YPH], p3 = y[YPP];
    v.norm = g[TT] * p0 * p0 + 2.0 * g[TP] * p0 * p3 + g[RR] * p1 * p1 + g[HH] * p2 * p2 +
             g[PP] * p3 * p3;
    const T sc = fabs(g[TT] * p0 * p0) + fabs(2.0 * g[TP] * p0 * p3) +
                 fabs(g[RR] * p1 * p1) + fabs(g[HH] * p2 * p2) +
                 fabs(g[PP] * p3 * p3);
    v.norm_rel = (sc > 0.0) ? v.norm / sc : v.norm;
    return v;
}
inline Inv invariants(double a, const double y[8]) { return invariants_t<double>(a, y); }

template <class T>
void to_cartesian_t(T a, T r, T th, T ph, T o[3]) {  // geodesic.h:256-262
    using std::cos; using std::sin; using std::sqrt;
    const T rr = sqrt(r * r + a * a);
    const T sn = sin(th);
    o[0] = rr * sn * cos(ph);
    o[1] = rr * sn * sin(ph);
    o[2] = r * cos(th);
}
inline void to_cartesian(double a, double r, double th, double ph, double o[3]) {
    to_cartesian_t<double>(a, r, th, ph, o);
}

// ---- spectrum (src/spectrum.cpp) -----------------------------------------
double planck(double lam, double T) {  // spectrum.cpp:13-22
    if (T <= 0.0 || lam <= 0.0) return 0.0;
    const double l = lam;
    const double l5 = l * l * l * l * l;
    const double x = kH * kC / (l * kKB * T);
    if (x > 700.0) return 0.0;
    return (2.0 * kH * kC * kC) / (l5 * std::expm1(x));
}

// Wyman, Sloan & Shirley (2013) multi-lobe fit of the CIE 1931 CMFs
// (spectrum.cpp:36-53).
double lobe(double x, double mu, double s1, double s2) {
    const double t = (x - mu) * ((x < mu) ? 1.0 / s1 : 1.0 / s2);
    return std::exp(-0.5 * t * t);
}
double cmf_x(double nm) {
    return 1.056 * lobe(nm, 599.8, 37.9, 31.0) + 0.362 * lobe(nm, 442.0, 16.0, 26.7) -
           0.065 * lobe(nm, 501.1, 20.4, 26.2);
}
double cmf_y(double nm) {
    return 0.821 * lobe(nm, 568.8, 46.9, 40.5) + 0.286 * lobe(nm, 530.9, 16.3, 31.1);
}
double cmf_z(double nm) {
    return 1.217 * lobe(nm, 437.0, 11.8, 36.0) + 0.681 * lobe(nm, 459.0, 26.0, 13.8);
}

void bb_exact(double T, double o[3]) {  // spectrum.cpp:55-73, 1 nm over 360-830 nm
    o[0] = o[1] = o[2] = 0.0;
    if (T <= 0.0) return;
    const double lo = 360.0, hi = 830.0, dl = 1.0;
    for (double nm = lo; nm <= hi; nm += dl) {
        const double w = planck(nm * 1e-9, T) * dl * 1e-9;
        o[0] += w * cmf_x(nm);
        o[1] += w * cmf_y(nm);
        o[2] += w * cmf_z(nm);
    }
}

void powerlaw(double alpha, double o[3]) {  // spectrum.cpp:77-92
    const double lambda0 = 550e-9;
    o[0] = o[1] = o[2] = 0.0;
    const double lo = 360.0, hi = 830.0, dl = 1.0;
    for (double nm = lo; nm <= hi; nm += dl) {
        const double lam = nm * 1e-9;
        const double Il = kC * std::pow(lam / lambda0, alpha) / (lam * lam);
        const double w = Il * dl * 1e-9;
        o[0] += w * cmf_x(nm);
        o[1] += w * cmf_y(nm);
        o[2] += w * cmf_z(nm);
    }
}

constexpr int kTabN = 4096;
constexpr double kLgMin = 1.6989700043360187;  // log10(50 K)
constexpr double kLgMax = 9.4771212547196626;  // log10(3e9 K)

struct BBTable {  // natural logs of X, Y, Z on a log10(T) grid (spectrum.cpp:97-114)
    std::vector<double> x, y, z;
    BBTable() : x(kTabN), y(kTabN), z(kTabN) {
        for (int i = 0; i < kTabN; ++i) {
            const double f = static_cast<double>(i) / (kTabN - 1);
            const double T = std::pow(10.0, kLgMin + f * (kLgMax - kLgMin));
            double v[3];
            bb_exact(T, v);
            x[i] = std::log(std::max(v[0], 1e-300));
            y[i] = std::log(std::max(v[1], 1e-300));
            z[i] = std::log(std::max(v[2], 1e-300));
        }
    }
};

const BBTable& bb_table() {
    static const BBTable t;
    return t;
}

void bb_xyz(double T, double o[3]) {  // spectrum.cpp:123-134
    const BBTable& t = bb_table();
    const double lg = std::log10(std::clamp(T, 50.0, 3.0e9));
    const double f = (lg - kLgMin) / (kLgMax - kLgMin) * (kTabN - 1);
    const int i = std::clamp(static_cast<int>(f), 0, kTabN - 2);
    const double u = f - i;
    o[0] = std::exp(t.x[i] * (1.0 - u) + t.x[i + 1] * u);
    o[1] = std::exp(t.y[i] * (1.0 - u) + t.y[i + 1] * u);
    o[2] = std::exp(t.z[i] * (1.0 - u) + t.z[i + 1] * u);
}

double bv_temperature(double bv) {  // spectrum.cpp:158-162 (Ballesteros 2012)
    const double x = std::clamp(bv, -0.4, 2.0);
    return 4600.0 * (1.0 / (0.92 * x + 1.7) + 1.0 / (0.92 * x + 0.62));
}

// ---- RNG: splitmix64 (scene.h:27-44) --------------------------------------
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (next() >> 11) * 0x1.0p-53; }
    double u(double lo, double hi) { return lo + (hi - lo) * u01(); }
    double gauss() {
        const double u1 = std::max(u01(), 1e-15), u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

// ---- star field (scene.h:60-223) -------------------------------------------
constexpr int kNLat = 360, kNLon = 720;

struct Sky {
    std::vector<double> x, y, z, T, irr;
    std::vector<std::vector<uint32_t>> bins;  // star indices per cell, ascending

    void bin_all() {
        bins.assign(static_cast<size_t>(kNLat) * kNLon, {});
        for (size_t i = 0; i < x.size(); ++i) {
            const double th = std::acos(std::clamp(z[i], -1.0, 1.0));
            const double ph = std::atan2(y[i], x[i]);
            int ti = static_cast<int>(th / kPi * kNLat);
            int pi = static_cast<int>((ph + kPi) / (2.0 * kPi) * kNLon);
            ti = std::clamp(ti, 0, kNLat - 1);
            pi = ((pi % kNLon) + kNLon) % kNLon;
            bins[static_cast<size_t>(ti) * kNLon + pi].push_back(static_cast<uint32_t>(i));
        }
    }
};

void generate_sky(int count, uint64_t seed, Sky& sk) {  // scene.h:62-107
    SplitMix rng(seed);
    const double ca = 0.8660254037844387, sa = 0.5;  // 30 degree tilt (scene.h:185-190)
    sk.x.resize(count); sk.y.resize(count); sk.z.resize(count);
    sk.T.resize(count); sk.irr.resize(count);
    for (int i = 0; i < count; ++i) {
        double d[3];
        if (rng.u01() < 0.45) {
            const double b = rng.gauss() * 0.12;
            const double l = rng.u(0.0, 2.0 * kPi);
            const double cb = std::cos(b);
            const double g[3] = {cb * std::cos(l), cb * std::sin(l), std::sin(b)};
            d[0] = g[0];
            d[1] = ca * g[1] - sa * g[2];
            d[2] = sa * g[1] + ca * g[2];
        } else {
            const double z = rng.u(-1.0, 1.0);
            const double ph = rng.u(0.0, 2.0 * kPi);
            const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
            d[0] = rr * std::cos(ph); d[1] = rr * std::sin(ph); d[2] = z;
        }
        sk.x[i] = d[0]; sk.y[i] = d[1]; sk.z[i] = d[2];
        const double k = 0.45, m_lo = -1.5, m_hi = 12.0;
        const double u = rng.u01();
        const double mag = m_lo + std::log10(u * (std::pow(10.0, k * (m_hi - m_lo)) - 1.0) + 1.0) / k;
        double bv = 0.55 + 0.42 * rng.gauss();
        if (rng.u01() < 0.10) bv = rng.u(1.2, 1.9);
        sk.T[i] = bv_temperature(std::clamp(bv, -0.35, 1.95));
        const double F0 = 3.64e-9 * 0.089 * 1e6 * 1e-6;
        sk.irr[i] = F0 * std::pow(10.0, -0.4 * mag);
    }
    sk.bin_all();
}

// The synthetic uniform sky of the 1024^2 frame (the library's
// bhx_uniform_sky; the reference has no such generator): uniform directions,
// m from N(<m) ~ 10^(0.45 m) on [-1.5, 8], B-V ~ N(0.6, 0.4), the reference's
// F0 and B-V -> T mapping; five splitmix64 draws per star.
void uniform_sky(int count, uint64_t seed, Sky& sk) {
    SplitMix rng(seed);
    sk.x.resize(count); sk.y.resize(count); sk.z.resize(count);
    sk.T.resize(count); sk.irr.resize(count);
    const double span = std::pow(10.0, 0.45 * (8.0 - (-1.5))) - 1.0;
    for (int i = 0; i < count; ++i) {
        const double z = 2.0 * rng.u01() - 1.0;
        const double ph = 2.0 * kPi * rng.u01();
        const double rr = std::sqrt(std::max(0.0, (1.0 - z) * (1.0 + z)));
        sk.x[i] = rr * std::cos(ph);
        sk.y[i] = rr * std::sin(ph);
        sk.z[i] = z;
        const double mag = -1.5 + std::log10(1.0 + span * rng.u01()) / 0.45;
        const double g = std::sqrt(-2.0 * std::log(1.0 - rng.u01()));
        sk.T[i] = bv_temperature(0.6 + 0.4 * g * std::cos(2.0 * kPi * rng.u01()));
        sk.irr[i] = 3.64e-9 * 0.089 * std::pow(10.0, -0.4 * mag);
    }
}

void milky_way(const double d[3], double o[3]) {  // scene.h:202-223
    const double ca = 0.8660254037844387, sa = 0.5;
    const double g[3] = {d[0], ca * d[1] + sa * d[2], -sa * d[1] + ca * d[2]};
    const double sin_b = std::clamp(g[2], -1.0, 1.0);
    const double band = std::exp(-std::fabs(sin_b) / 0.055);
    const double l = std::atan2(g[1], g[0]);
    const double bulge = 0.55 + 0.45 * std::exp(-(l * l) / (2.0 * 0.45 * 0.45));
    const double lane = 0.65 + 0.35 * std::sin(l * 7.3 + sin_b * 41.0) * std::sin(l * 3.1 - sin_b * 17.0);
    const double mag = 21.5;
    const double F0 = 3.24e-10;
    const double per_sr = 4.2545170296152206e10;
    const double L = F0 * std::pow(10.0, -0.4 * mag) * per_sr;
    const double amp = L * band * bulge * std::max(0.0, lane);
    double c[3];
    bb_xyz(4800.0, c);
    const double inv = (c[1] > 0.0) ? 1.0 / c[1] : 0.0;
    o[0] = amp * c[0] * inv;
    o[1] = amp;
    o[2] = amp * c[2] * inv;
}

void sky_radiance(const Sky& sk, const double d[3], double sigma, double o[3]) {  // scene.h:111-158
    o[0] = o[1] = o[2] = 0.0;
    const double cutoff = 3.0 * sigma;
    const double inv2s2 = 1.0 / (2.0 * sigma * sigma);
    const double norm = 1.0 / (2.0 * kPi * sigma * sigma);
    const double theta = std::acos(std::clamp(d[2], -1.0, 1.0));
    const double phi = std::atan2(d[1], d[0]);
    const int it = static_cast<int>(theta / kPi * kNLat);
    const double dlat = kPi / kNLat;
    const int span_t = std::max(1, static_cast<int>(std::ceil(cutoff / dlat)));
    const double sn = std::max(std::sin(theta), 1e-6);
    const double dlon = 2.0 * kPi / kNLon;
    const int span_p = std::max(1, static_cast<int>(std::ceil(cutoff / (sn * dlon))));
    for (int da = -span_t; da <= span_t; ++da) {
        const int ti = it + da;
        if (ti < 0 || ti >= kNLat) continue;
        const int sp = (span_p * 2 + 1 >= kNLon) ? kNLon / 2 : span_p;
        for (int db = -sp; db <= sp; ++db) {
            int pi = static_cast<int>((phi + kPi) / (2.0 * kPi) * kNLon) + db;
            pi = ((pi % kNLon) + kNLon) % kNLon;
            for (uint32_t idx : sk.bins[static_cast<size_t>(ti) * kNLon + pi]) {
                const double dot =
                    std::clamp(d[0] * sk.x[idx] + d[1] * sk.y[idx] + d[2] * sk.z[idx], -1.0, 1.0);
                const double ang = (dot > 0.9999995) ? std::sqrt(std::max(0.0, 2.0 * (1.0 - dot)))
                                                     : std::acos(dot);
                if (ang > cutoff) continue;
                const double w = norm * std::exp(-ang * ang * inv2s2) * sk.irr[idx];
                double c[3];
                bb_xyz(sk.T[idx], c);
                const double inv = (c[1] > 0.0) ? 1.0 / c[1] : 0.0;
                o[0] += w * c[0] * inv;
                o[1] += w;
                o[2] += w * c[2] * inv;
            }
        }
    }
    double mw[3];
    milky_way(d, mw);
    o[0] += mw[0]; o[1] += mw[1]; o[2] += mw[2];
}

// ---- Page-Thorne / Novikov-Thorne disc (src/disc.h) ------------------------
double pt_W(double a, double r, double r_isco) {  // disc.h:46-64
    const double x = std::sqrt(r);
    const double x0 = std::sqrt(r_isco);
    const double ac = std::acos(std::clamp(a, -1.0, 1.0));
    const double x1 = 2.0 * std::cos(ac / 3.0 - kPi / 3.0);
    const double x2 = 2.0 * std::cos(ac / 3.0 + kPi / 3.0);
    const double x3 = -2.0 * std::cos(ac / 3.0);
    auto term = [&](double xi, double xj, double xk) {
        if (std::fabs(xi) < 1e-13) return 0.0;
        return 3.0 * (xi - a) * (xi - a) / (xi * (xi - xj) * (xi - xk)) *
               std::log((x - xi) / (x0 - xi));
    };
    return x - x0 - 1.5 * a * std::log(x / x0) - term(x1, x2, x3) - term(x2, x1, x3) -
           term(x3, x1, x2);
}

struct Disc {  // disc.h:66-151
    double a, r_in, r_out, eta, mdot, rg;
    int sg;
    Disc(double M, double spin, double rout, double edd, int sense) {
        a = std::clamp(spin, -0.9999, 0.9999);
        sg = sense;
        r_in = isco(std::clamp(spin, -0.9999, 0.9999), sense);
        r_out = std::max(rout, isco(std::clamp(spin, -0.9999, 0.9999), sense) * 1.02);
        eta = 1.0 - circ_energy(a, isco(a, sg), sg);
        mdot = edd * eddington_mdot(M, eta);
        rg = rg_metres(M);
    }
    bool holds(double r) const { return r >= r_in && r <= r_out; }
    double omega(double r) const {
        const double r32 = r * std::sqrt(r);
        return sg / (r32 + sg * a);
    }
    double omega_dr(double r) const {
        const double r32 = r * std::sqrt(r);
        const double d = r32 + sg * a;
        return -sg * 1.5 * std::sqrt(r) / (d * d);
    }
    double flux_dimless(double r) const {
        if (r <= r_in) return 0.0;
        const double W = pt_W(a, r, r_in);
        const double Om = omega(r);
        const double dOm = omega_dr(r);
        const double E = circ_energy(a, r, sg);
        const double L = circ_angmom(a, r, sg);
        const double den = (E - Om * L);
        if (!(den > 0.0)) return 0.0;
        const double F = (-dOm) / (4.0 * kPi * r * den * den) * W;
        return (F > 0.0 && std::isfinite(F)) ? F : 0.0;
    }
    double flux(double r) const {
        if (r <= r_in || r > r_out) return 0.0;
        return flux_dimless(r) * mdot * kC * kC / (rg * rg);
    }
    double temperature(double r) const {
        const double F = flux(r);
        if (F <= 0.0) return 0.0;
        return std::pow(F / kSigmaSB, 0.25);
    }
};

// ---- jets (src/jet.h:89-248, 314) -----------------------------------------
struct JetModel {
    double z_base, z_top, R_base, k, gamma_max, z_sat, alpha, u_peak, u_wid, spine, spine_wid;
    double zeta, Omega_F, j_nu0, r_bound;
    // conical jet (bhx_scene.jet_model = 1), this package's own model: no
    // reference counterpart; the same formulas as the device's cone_shape /
    // cone_boost, written with the C library's functions.  Profile (r0/r)^2
    // x taper x end ramps, all three with the quintic step (the taper in
    // |cos theta|, from the cone's edge to 0.8 of its opening); sampled by 8
    // sub-intervals of 4-point Gauss-Legendre on the steps whose dense output
    // can reach the cone: |theta(s) - theta(0)| <= |c2| + |c3|/4 + 4|c4|/27 +
    // |c5|/16 over the step.
    int model = 0;
    double cone_angle = 0.0, cone_r0 = 0.0, cone_r1 = 0.0;
    static double smooth_c2(double t) {
        t = std::clamp(t, 0.0, 1.0);
        return t * t * t * (10.0 + t * (6.0 * t - 15.0));
    }
    double cone_shape(double r, double cs) const {
        if (r <= cone_r0 || r >= cone_r1) return 0.0;
        const double c0 = std::cos(cone_angle), c1 = std::cos(0.8 * cone_angle);
        const double mu = std::fabs(cs);
        if (mu <= c0) return 0.0;   // outside the cone
        const double ta = smooth_c2((mu - c0) / (c1 - c0));
        const double q = cone_r0 / r;
        const double ramp = smooth_c2((r - cone_r0) / (0.5 * cone_r0)) * smooth_c2((cone_r1 - r) / (0.15 * cone_r1));
        return q * q * ta * ramp;
    }
    // the step's dense output (theta coefficients c1..c5) can reach the cone
    bool cone_gate(double c1, double c2, double c3, double c4, double c5) const {
        const double reach = std::fabs(c2) + std::fabs(c3) / 4.0 + 4.0 * std::fabs(c4) / 27.0 + std::fabs(c5) / 16.0;
        const double psi0 = std::min(c1, 3.14159265358979323846 - c1);
        return psi0 - reach < cone_angle;
    }
    void cone_velocity(const Pt& p, double u[4]) const {
        const double G = std::max(1.0, gamma_max);
        const double beta = std::sqrt(std::max(0.0, 1.0 - 1.0 / (G * G)));
        const double lapse = std::sqrt(p.S * p.D / p.A);
        const double omega = 2.0 * p.a * p.r / p.A;
        u[0] = G / lapse;
        u[1] = G * beta * std::sqrt(p.D / p.S);
        u[2] = 0.0;
        u[3] = G * omega / lapse;
    }
    double radius(double az) const { return R_base * std::pow(std::max(az, 1e-6) / z_base, k); }
    double gamma_at(double az) const {
        const double t = std::min(1.0, std::pow(std::max(az, 1e-6) / z_sat, k));
        return 1.0 + (gamma_max - 1.0) * t;
    }
    double speed_at(double az) const {
        const double G = gamma_at(az);
        return std::sqrt(std::max(0.0, 1.0 - 1.0 / (G * G)));
    }
    double profile(double u) const {
        const double s = (u - u_peak) / u_wid;
        const double t = u / spine_wid;
        return std::exp(-s * s) + spine * std::exp(-t * t);
    }
    static double smooth01(double t) {
        t = std::clamp(t, 0.0, 1.0);
        return t * t * (3.0 - 2.0 * t);
    }
    double shape(double z, double rho) const {
        const double az = std::fabs(z);
        if (az <= z_base || az >= z_top) return 0.0;
        const double R = radius(az);
        const double f = profile(rho / R);
        if (f < 1e-9) return 0.0;
        const double g = std::pow(gamma_at(az) / gamma_at(z_base), -(2.0 + alpha)) *
                         std::pow(R / R_base, -(3.0 + alpha)) * std::pow(az / z_base, zeta);
        const double ramp = smooth01((az - z_base) / (0.5 * z_base)) *
                            smooth01((z_top - az) / (0.15 * z_top));
        return f * g * ramp;
    }
    void four_velocity(const Pt& p, double u[4]) const {
        double e[4][4];
        zamo(p, e);
        const double lapse = std::sqrt(p.S * p.D / p.A);
        const double omega = 2.0 * p.a * p.r / p.A;
        const double sqrt_gpp = p.sn * std::sqrt(p.A / p.S);
        const double z = p.r * p.cs;
        const double v = speed_at(std::fabs(z));
        double v_phi = (Omega_F - omega) * sqrt_gpp / std::max(lapse, 1e-12);
        v_phi = std::clamp(v_phi, -0.5 * v, 0.5 * v);
        const double v_pol = std::sqrt(std::max(0.0, v * v - v_phi * v_phi));
        const double sgn = (p.cs >= 0.0) ? 1.0 : -1.0;
        const double d_r = sgn * p.cs;
        const double d_th = -sgn * p.sn;
        const double G = 1.0 / std::sqrt(std::max(1e-12, 1.0 - v * v));
        for (int mu = 0; mu < 4; ++mu)
            u[mu] = G * (e[0][mu] + v_pol * (d_r * e[1][mu] + d_th * e[2][mu]) + v_phi * e[3][mu]);
    }
    double boost(double g) const { return std::pow(g, 2.0 + alpha); }
};

constexpr double kGLx[4] = {0.06943184420297371, 0.33000947820757187, 0.66999052179242813,
                            0.93056815579702629};
constexpr double kGLw[4] = {0.17392742256872693, 0.32607257743127307, 0.32607257743127307,
                            0.17392742256872693};

double seg_dist_origin(const double p0[3], const double p1[3]) {  // render.cpp:55-62
    const double d[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double t = 0.0;
    if (dd > 0.0) t = std::clamp(-(p0[0] * d[0] + p0[1] * d[1] + p0[2] * d[2]) / dd, 0.0, 1.0);
    const double q[3] = {p0[0] + t * d[0], p0[1] + t * d[1], p0[2] + t * d[2]};
    return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
}

double redshift(const double y[8], const double u[4]) {  // render.cpp:31-34
    const double ku = y[YPT] * u[0] + y[YPR] * u[1] + y[YPH] * u[2] + y[YPP] * u[3];
    return (ku > 1e-12) ? 1.0 / ku : 0.0;
}

void cart_velocity(double a, const double y[8], const double dy[8], double v[3]) {  // render.cpp:15-26
    const double r = y[YR], th = y[YH], ph = y[YP];
    const double R = std::sqrt(r * r + a * a);
    const double sn = std::sin(th), cs = std::cos(th);
    const double sp = std::sin(ph), cp = std::cos(ph);
    const double dr = dy[YR], dth = dy[YH], dph = dy[YP];
    const double dR = (R > 0.0) ? r / R * dr : dr;
    v[0] = dR * sn * cp + R * cs * dth * cp - R * sn * sp * dph;
    v[1] = dR * sn * sp + R * cs * dth * sp + R * sn * cp * dph;
    v[2] = dr * cs - r * sn * dth;
}

// ---- scene constants -------------------------------------------------------
struct Scene {
    bhx_scene c;
    double a, r_h, r_stop, psf_sigma, aspect, jet_rgb[3];
    double tet[4][4];
    bool has_disc;
    Disc* disc = nullptr;
    JetModel jet{};
    const Sky* sky = nullptr;
    Sky own_sky;

    explicit Scene(const bhx_scene* s, const oracle_catalogue* cat) : c(*s) {
        a = c.spin;
        r_h = horizon_plus(a);
        r_stop = r_h * 1.0 + 1e-4 + (r_h - horizon_minus(a)) * 1e-3;  // render.cpp:72-73
        has_disc = c.disc_enabled != 0;
        if (has_disc) disc = new Disc(c.M_kg, c.spin, c.disc_r_out, c.eddington, c.disc_sense);
        if (c.stars_enabled) {
            if (cat && cat->n > 0) {
                own_sky.x.assign(cat->x, cat->x + cat->n);
                own_sky.y.assign(cat->y, cat->y + cat->n);
                own_sky.z.assign(cat->z, cat->z + cat->n);
                own_sky.T.assign(cat->temperature, cat->temperature + cat->n);
                own_sky.irr.assign(cat->irradiance, cat->irradiance + cat->n);
                own_sky.bin_all();
            } else {
                generate_sky(c.star_count, c.seed, own_sky);
            }
            sky = &own_sky;
        }
        zamo(point(c.spin, c.cam_r, c.cam_theta), tet);
        aspect = static_cast<double>(c.width) / c.height;
        const double px_ang = c.cam_fov / c.width;
        psf_sigma = 0.55 * px_ang;
        jet_rgb[0] = jet_rgb[1] = jet_rgb[2] = 0.0;
        if (c.jet_enabled) powerlaw(c.jet_alpha, jet_rgb);
        jet = JetModel{c.jet_z_base, c.jet_z_top, c.jet_R_base, c.jet_k, c.jet_gamma_max,
                       c.jet_z_sat, c.jet_alpha, c.jet_u_peak, c.jet_u_wid, c.jet_spine,
                       c.jet_spine_wid, c.jet_zeta, c.jet_Omega_F, c.jet_j_nu0, c.jet_r_bound};
        jet.model = c.jet_model;
        jet.cone_angle = c.jet_cone_angle;
        jet.cone_r0 = c.jet_cone_r0;
        jet.cone_r1 = c.jet_cone_r1;
    }
    ~Scene() { delete disc; }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    // Camera::ray_direction (scene.h:356-384)
    void ray_dir(double sx, double sy, double o[3]) const {
        const double tanf = std::tan(0.5 * c.cam_fov);
        double x = sx * tanf;
        double y = sy * tanf / aspect;
        double z = 1.0;
        if (c.cam_roll != 0.0) {
            const double cr = std::cos(c.cam_roll), sr = std::sin(c.cam_roll);
            const double xr = cr * x - sr * y, yr = sr * x + cr * y;
            x = xr; y = yr;
        }
        const double cp = std::cos(c.cam_pitch), sp = std::sin(c.cam_pitch);
        const double y2 = cp * y - sp * z;
        const double z2 = sp * y + cp * z;
        const double cy = std::cos(c.cam_yaw), sy2 = std::sin(c.cam_yaw);
        const double x2 = cy * x + sy2 * z2;
        const double z3 = -sy2 * x + cy * z2;
        const double v[3] = {-z3, -y2, x2};
        const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        o[0] = v[0] / n; o[1] = v[1] / n; o[2] = v[2] / n;
    }

    // Direction of sample k of pixel (px, py), with the reference's jitter
    // stream (render.cpp:437-451).
    void sample_dir(int px, int py, int k, double o[3]) const {
        const int S = std::max(1, c.sqrt_spp);
        const double inv_S = 1.0 / S;
        SplitMix rng(c.seed ^ (static_cast<uint64_t>(py) << 32) ^
                     (static_cast<uint64_t>(px) * 0x9E3779B97F4A7C15ull));
        double jx = 0.0, jy = 0.0;
        for (int s = 0; s <= k; ++s) {
            const int sy = s / S, sx = s % S;
            jx = (sx + rng.u01()) * inv_S;
            jy = (sy + rng.u01()) * inv_S;
        }
        const double u = (2.0 * (px + jx) / c.width - 1.0);
        const double v = (1.0 - 2.0 * (py + jy) / c.height);
        ray_dir(u, v, o);
    }

    // NeutronStar::surface_four_velocity (neutronstar.h:90-95)
    void ns_u(double th, double u[4]) const {
        const double sn = std::sin(th);
        const double R = c.ns_R, Om = c.ns_Omega;
        const double norm2 = 1.0 - 2.0 / R - Om * Om * R * R * sn * sn;
        const double ut = (norm2 > 1e-12) ? 1.0 / std::sqrt(norm2) : 1.0;
        u[0] = ut; u[1] = 0.0; u[2] = 0.0; u[3] = Om * ut;
    }
    // NeutronStar::temperature_at (neutronstar.h:100-125)
    double ns_T(double th, double ph, double t_emit) const {
        if (!c.ns_caps) return c.ns_T_eff;
        const double phase = c.ns_cap_phase0 + c.ns_Omega * t_emit;
        const double m[3] = {std::sin(c.ns_cap_tilt) * std::cos(phase),
                             std::sin(c.ns_cap_tilt) * std::sin(phase), std::cos(c.ns_cap_tilt)};
        const double n[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
        const double dot = n[0] * m[0] + n[1] * m[1] + n[2] * m[2];
        double ang = std::acos(std::clamp(dot, -1.0, 1.0));
        if (c.ns_two_caps) ang = std::min(ang, kPi - ang);
        if (ang <= c.ns_cap_radius) {
            const double edge = 0.9 * c.ns_cap_radius;
            if (ang <= edge) return c.ns_cap_T;
            const double u = (ang - edge) / (c.ns_cap_radius - edge);
            const double w = 0.5 * (1.0 + std::cos(kPi * u));
            return c.ns_T_eff + (c.ns_cap_T - c.ns_T_eff) * w;
        }
        return c.ns_T_eff;
    }
    // HotSpot::enhancement (scene.h:325-334)
    double hotspot(double r_hit, double phi_hit, double t_emit) const {
        if (!c.hs_enabled) return 1.0;
        const double phi_s = c.hs_phi0 + c.hs_omega * t_emit;
        const double dx = r_hit * std::cos(phi_hit) - c.hs_r * std::cos(phi_s);
        const double dy = r_hit * std::sin(phi_hit) - c.hs_r * std::sin(phi_s);
        const double d2 = dx * dx + dy * dy;
        return 1.0 + c.hs_contrast * std::exp(-0.5 * d2 / (c.hs_sigma * c.hs_sigma));
    }
};

struct RayResult {
    int cls = BHX_EXHAUSTED;
    double y[8];
    double jet = 0.0;
    int steps = 0;
    int attempts = 0;
    double norm = 0.0, carter = 0.0;
    long long rhs = 0;
    double g = 0.0, dir[3] = {0, 0, 0}, xyz[3] = {0, 0, 0};
};

// trace() (render.cpp:69-260) from an initial photon state.
void trace_state(const Scene& sc, const double y0[8], RayResult& out) {
    const double a = sc.a;
    const bhx_scene& c = sc.c;
    double y[8];
    std::memcpy(y, y0, sizeof y);
    const Inv iv0 = invariants(a, y);
    const double Q0 = std::fabs(iv0.Q) + 1e-12;
    DP5 st{a, c.rtol, c.atol, {}, false, 1e-4, 0};
    st.reset();
    double h = 0.02 * std::max(1.0, c.cam_r);
    out.jet = 0.0;
    std::memcpy(out.y, y, sizeof y);
    double p0[3];
    to_cartesian(a, y[YR], y[YH], y[YP], p0);

    // the conical jet (this package's own model) has a sampler of its own so
    // that the reference's path stays the reference's code
    auto cone_sample = [&](const Seg& seg, double h_step, double s_end) {
        const int nsub = 8;
        double sum = 0.0;
        for (int k = 0; k < nsub; ++k)
            for (int q = 0; q < 4; ++q) {
                double ys[8];
                seg.at((k + kGLx[q]) / nsub * s_end, ys);
                const double sh = sc.jet.cone_shape(ys[YR], std::cos(ys[YH]));
                if (sh <= 0.0) continue;
                double u[4];
                sc.jet.cone_velocity(point(a, ys[YR], ys[YH]), u);
                const double gf = redshift(ys, u);
                if (gf <= 0.0) continue;
                sum += kGLw[q] * sh * sc.jet.boost(gf);
            }
        out.jet += c.jet_j_nu0 * sum * h_step * s_end / nsub;
    };
    auto jet_sample = [&](const Seg& seg, double h_step, double s_end) {
        if (!c.jet_enabled || s_end <= 0.0) return;
        if (sc.jet.model == 1) { cone_sample(seg, h_step, s_end); return; }
        double sum = 0.0;
        for (int q = 0; q < 4; ++q) {
            double ys[8];
            seg.at(kGLx[q] * s_end, ys);
            const double z = ys[YR] * std::cos(ys[YH]);
            const double rho = std::sqrt(ys[YR] * ys[YR] + a * a) * std::sin(ys[YH]);
            const double sh = sc.jet.shape(z, rho);
            if (sh <= 0.0) continue;
            double u[4];
            sc.jet.four_velocity(point(a, ys[YR], ys[YH]), u);
            const double gf = redshift(ys, u);
            if (gf <= 0.0) continue;
            sum += kGLw[q] * sh * sc.jet.boost(gf);
        }
        out.jet += c.jet_j_nu0 * sum * h_step * s_end;
    };

    auto done = [&](int cls) {
        out.cls = cls;
        out.rhs = st.rhs_calls;
    };

    for (int step = 0; step < c.max_steps; ++step) {
        double yn[8];
        Seg seg;
        const double hmax = 0.05 * std::max(1.0, y[YR] - sc.r_h) + 0.01;
        if (h > hmax) h = hmax;
        int tries = 0;
        ++out.attempts;
        while (!st.attempt(y, h, yn, seg)) {
            if (++tries > 40 || std::fabs(h) < 1e-14) { done(BHX_HORIZON); return; }
            ++out.attempts;
        }
        ++out.steps;
        {
            const Inv iv = invariants(a, yn);
            out.norm = std::max(out.norm, std::fabs(iv.norm_rel));
            out.carter = std::max(out.carter, std::fabs(iv.Q - iv0.Q) / Q0);
        }
        double p1[3];
        to_cartesian(a, yn[YR], yn[YH], yn[YP], p1);
        bool near_jet = c.jet_enabled && seg_dist_origin(p0, p1) < c.jet_r_bound;
        if (near_jet && sc.jet.model == 1)
            near_jet = sc.jet.cone_gate(seg.c1[YH], seg.c2[YH], seg.c3[YH], seg.c4[YH], seg.c5[YH]);

        if (c.ns_enabled && yn[YR] <= c.ns_R) {  // render.cpp:158-169
            double lo = 0.0, hi = 1.0;
            for (int k = 0; k < 60; ++k) {
                const double mid = 0.5 * (lo + hi);
                if (seg.comp(mid, YR) <= c.ns_R) hi = mid; else lo = mid;
            }
            const double s_end = 0.5 * (lo + hi);
            if (near_jet) jet_sample(seg, seg.h, s_end);
            seg.at(s_end, out.y);
            done(BHX_STAR);
            return;
        }
        if (yn[YR] <= sc.r_stop) {  // render.cpp:172-177
            if (near_jet) jet_sample(seg, seg.h, 1.0);
            std::memcpy(out.y, yn, sizeof yn);
            done(BHX_HORIZON);
            return;
        }
        if (sc.has_disc) {  // render.cpp:180-199
            const double c0 = std::cos(y[YH]);
            const double c1 = std::cos(yn[YH]);
            if (c0 * c1 < 0.0) {
                double lo = 0.0, hi = 1.0;
                for (int k = 0; k < 60; ++k) {
                    const double mid = 0.5 * (lo + hi);
                    const double cm = std::cos(seg.comp(mid, YH));
                    if (cm * c0 <= 0.0) hi = mid; else lo = mid;
                }
                const double s_end = 0.5 * (lo + hi);
                double yc[8];
                seg.at(s_end, yc);
                if (sc.disc->holds(yc[YR])) {
                    if (near_jet) jet_sample(seg, seg.h, s_end);
                    std::memcpy(out.y, yc, sizeof yc);
                    done(BHX_DISC);
                    return;
                }
            }
        }
        if (c.sun_enabled) {  // render.cpp:202-241
            const double* C = c.sun_centre;
            const double sv[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            const double seg_len = std::sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
            const double dc = std::sqrt((p0[0] - C[0]) * (p0[0] - C[0]) + (p0[1] - C[1]) * (p0[1] - C[1]) +
                                        (p0[2] - C[2]) * (p0[2] - C[2]));
            if (dc <= seg_len + c.sun_radius * 1.5) {
                auto f = [&](double s) {
                    double ys[8], q[3];
                    seg.at(s, ys);
                    to_cartesian(a, ys[YR], ys[YH], ys[YP], q);
                    const double dx = q[0] - C[0], dy2 = q[1] - C[1], dz = q[2] - C[2];
                    return dx * dx + dy2 * dy2 + dz * dz - c.sun_radius * c.sun_radius;
                };
                const int kProbe = 12;
                if (f(0.0) <= 0.0) {
                    std::memcpy(out.y, y, sizeof y);
                    done(BHX_SUN);
                    return;
                }
                double prev = 0.0;
                for (int k = 1; k <= kProbe; ++k) {
                    const double s = static_cast<double>(k) / kProbe;
                    if (f(s) <= 0.0) {
                        double lo = prev, hi2 = s;
                        for (int b = 0; b < 50; ++b) {
                            const double mid = 0.5 * (lo + hi2);
                            if (f(mid) <= 0.0) hi2 = mid; else lo = mid;
                        }
                        const double s_end = 0.5 * (lo + hi2);
                        if (near_jet) jet_sample(seg, seg.h, s_end);
                        seg.at(s_end, out.y);
                        done(BHX_SUN);
                        return;
                    }
                    prev = s;
                }
            }
        }
        if (near_jet) jet_sample(seg, seg.h, 1.0);
        if (yn[YR] > c.r_escape) {  // render.cpp:247-251
            std::memcpy(out.y, yn, sizeof yn);
            done(BHX_SKY);
            return;
        }
        std::memcpy(y, yn, sizeof y);
        p0[0] = p1[0]; p0[1] = p1[1]; p0[2] = p1[2];
    }
    std::memcpy(out.y, y, sizeof y);
    done(BHX_EXHAUSTED);
}

// shade_hit() + shade() (render.cpp:265-393); also records g and the sky
// direction, which the reference computes but does not return.
void shade_ray(const Scene& sc, RayResult& r, bhx_stats& st) {
    const double a = sc.a;
    const bhx_scene& c = sc.c;
    double col[3] = {0.0, 0.0, 0.0};
    r.g = 0.0;
    r.dir[0] = r.dir[1] = r.dir[2] = 0.0;
    switch (r.cls) {
        case BHX_HORIZON:
            ++st.captured;
            break;
        case BHX_DISC: {
            ++st.disc_hits;
            const double rr = r.y[YR];
            double T = sc.disc->temperature(rr);
            if (T <= 0.0) break;
            if (c.hs_enabled) {
                const double t_emit = c.observer_time + r.y[YT];
                const double f = sc.hotspot(rr, r.y[YP], t_emit);
                T *= std::pow(f, 0.25);
            }
            double u[4];
            kepler_u(sc.disc->a, rr, sc.disc->sg, u);
            const double g = redshift(r.y, u);
            r.g = g;
            if (g <= 0.0) break;
            bb_xyz(g * T, col);
            break;
        }
        case BHX_SUN: {
            ++st.sun_hits;
            double gl[5];
            cov_metric(point(a, r.y[YR], r.y[YH]), gl);
            if (gl[TT] >= 0.0) break;
            const double u[4] = {1.0 / std::sqrt(-gl[TT]), 0.0, 0.0, 0.0};
            const double g = redshift(r.y, u);
            r.g = g;
            if (g <= 0.0) break;
            double p[3];
            to_cartesian(a, r.y[YR], r.y[YH], r.y[YP], p);
            double d[8], v[3];
            rhs(a, r.y, d);
            cart_velocity(a, r.y, d, v);
            const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (vn > 0.0) { v[0] /= vn; v[1] /= vn; v[2] /= vn; }
            // SunBody::limb_darkening (scene.h:266-276)
            double limb = 1.0;
            {
                double n[3] = {p[0] - c.sun_centre[0], p[1] - c.sun_centre[1], p[2] - c.sun_centre[2]};
                const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                if (len > 0.0) {
                    n[0] /= len; n[1] /= len; n[2] /= len;
                    const double mu = std::clamp(-(n[0] * v[0] + n[1] * v[1] + n[2] * v[2]), 0.0, 1.0);
                    limb = (2.0 + 3.0 * mu) / 5.0;
                }
            }
            const double self_z = kG * kMsun / (kRsun * kC * kC);  // scene.h:239-240
            const double T_obs = g * c.sun_T_eff / (1.0 + self_z);
            bb_xyz(T_obs, col);
            col[0] *= limb; col[1] *= limb; col[2] *= limb;
            break;
        }
        case BHX_STAR: {
            ++st.star_hits;
            const double rr = r.y[YR], th = r.y[YH], ph = r.y[YP];
            double u[4];
            sc.ns_u(th, u);
            const double ku = r.y[YPT] * u[0] + r.y[YPP] * u[3];
            if (!(ku > 1e-12)) break;
            const double g = 1.0 / ku;
            r.g = g;
            const double t_emit = c.observer_time + r.y[YT];
            const double T_local = sc.ns_T(th, ph, t_emit);
            const double f = 1.0 - 2.0 / rr;
            const double mu = (f > 0.0) ? std::clamp(-r.y[YPR] * std::sqrt(f) / ku, 0.0, 1.0) : 1.0;
            const double limb = (2.0 + 3.0 * mu) / 5.0;
            bb_xyz(g * T_local, col);
            col[0] *= limb; col[1] *= limb; col[2] *= limb;
            break;
        }
        case BHX_SKY: {
            ++st.escaped;
            double d[8], v[3];
            rhs(a, r.y, d);
            cart_velocity(a, r.y, d, v);
            const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (vn <= 0.0) break;
            v[0] /= vn; v[1] /= vn; v[2] /= vn;
            r.dir[0] = v[0]; r.dir[1] = v[1]; r.dir[2] = v[2];
            if (!sc.sky) break;
            sky_radiance(*sc.sky, v, sc.psf_sigma, col);
            break;
        }
        default:
            ++st.exhausted;
            break;
    }
    if (r.jet > 0.0) {
        ++st.jet_rays;
        col[0] += sc.jet_rgb[0] * r.jet;
        col[1] += sc.jet_rgb[1] * r.jet;
        col[2] += sc.jet_rgb[2] * r.jet;
    }
    r.xyz[0] = col[0]; r.xyz[1] = col[1]; r.xyz[2] = col[2];
}

void trace_one(const Scene& sc, int64_t ray, RayResult& r, bhx_stats& st) {
    const int S = std::max(1, sc.c.sqrt_spp);
    const int64_t pix = ray / (S * S);
    const int k = static_cast<int>(ray % (S * S));
    const int px = static_cast<int>(pix % sc.c.width), py = static_cast<int>(pix / sc.c.width);
    double dir[3], y0[8];
    sc.sample_dir(px, py, k, dir);
    photon(sc.a, sc.c.cam_r, sc.c.cam_theta, sc.c.cam_phi, sc.tet, dir, y0);
    trace_state(sc, y0, r);
    ++st.rays;
    st.steps += r.steps;
    st.attempts += r.attempts;
    st.rhs_evals += r.rhs;
    st.max_norm_error = std::max(st.max_norm_error, r.norm);
    st.max_carter_drift = std::max(st.max_carter_drift, r.carter);
    shade_ray(sc, r, st);
}

void merge(bhx_stats& d, const bhx_stats& s) {
    d.rays += s.rays; d.steps += s.steps; d.captured += s.captured; d.disc_hits += s.disc_hits;
    d.sun_hits += s.sun_hits; d.star_hits += s.star_hits; d.jet_rays += s.jet_rays;
    d.escaped += s.escaped; d.exhausted += s.exhausted; d.total_flux += s.total_flux;
    d.attempts += s.attempts; d.rhs_evals += s.rhs_evals;
    d.max_norm_error = std::max(d.max_norm_error, s.max_norm_error);
    d.max_carter_drift = std::max(d.max_carter_drift, s.max_carter_drift);
}

int n_threads(int req) {  // default: all cores, at most 16 (one GPU host's CPU share)
    if (req > 0) return req;
    const int n = static_cast<int>(std::thread::hardware_concurrency());
    return n > 0 ? std::min(n, 16) : 1;
}

}  // namespace

namespace {

// ---- algorithmic operation counts -------------------------------------------
// A scalar that counts its arithmetic: +, -, *, / and sqrt are one operation
// each, and so is each transcendental (sin, cos, pow); comparisons, fabs,
// max / min / clamp and negation are free.  Instantiating the templates
// above with it counts the reference's expressions as written.
struct Flop {
    double v = 0.0;
    Flop() = default;
    Flop(double x) : v(x) {}  // NOLINT: constants promote
};
thread_local long long g_flops = 0;
inline Flop op(double v) { ++g_flops; return Flop(v); }
inline Flop operator+(Flop a, Flop b) { return op(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { return op(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { return op(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { return op(a.v / b.v); }
inline Flop operator+(Flop a, double b) { return op(a.v + b); }
inline Flop operator-(Flop a, double b) { return op(a.v - b); }
inline Flop operator*(Flop a, double b) { return op(a.v * b); }
inline Flop operator/(Flop a, double b) { return op(a.v / b); }
inline Flop operator+(double a, Flop b) { return op(a + b.v); }
inline Flop operator-(double a, Flop b) { return op(a - b.v); }
inline Flop operator*(double a, Flop b) { return op(a * b.v); }
inline Flop operator/(double a, Flop b) { return op(a / b.v); }
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline Flop& operator+=(Flop& a, Flop b) { a = a + b; return a; }
inline Flop& operator*=(Flop& a, Flop b) { a = a * b; return a; }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator<(Flop a, double b) { return a.v < b; }
inline bool operator<(double a, Flop b) { return a < b.v; }
inline bool operator>(Flop a, double b) { return a.v > b; }
inline bool operator<=(Flop a, double b) { return a.v <= b; }
inline Flop sqrt(Flop a) { return op(std::sqrt(a.v)); }
inline Flop sin(Flop a) { return op(std::sin(a.v)); }
inline Flop cos(Flop a) { return op(std::cos(a.v)); }
inline Flop pow(Flop a, double b) { return op(std::pow(a.v, b)); }
inline Flop fabs(Flop a) { return Flop(std::fabs(a.v)); }
inline bool finite_t(Flop x) { return std::isfinite(x.v); }

// A ray state well inside the integration region (r = 12 M, equatorial
// photon of the default camera's kind), for counting.
void flop_state(double y[8]) {
    const double s[8] = {0.0, 12.0, 1.3, 0.4, 0.998, -0.31, 2.1, -3.7};
    for (int i = 0; i < 8; ++i) y[i] = s[i];
}

}  // namespace

extern "C" {

// Algorithmic FP64 operations of the reference's path, counted by
// instantiating the restatement with Flop (out: per RHS evaluation; per DP5
// attempt excluding its 6 RHS, accepted (incl. dense output and controller)
// and rejected; per accepted step of the trace loop outside the attempt:
// invariants, pseudo-Cartesian end point, the hmax clamp, the drift maxima and
// the disc's sign test (render.cpp:130-131, 140-145, 150, 181-183)).
void oracle_flop_counts(int64_t* out) {
    double y0[8];
    flop_state(y0);
    Flop y[8], d[8];
    for (int i = 0; i < 8; ++i) y[i] = y0[i];
    const Flop a = 0.9;
    g_flops = 0;
    rhs_t<Flop>(a, y, d);
    const long long per_rhs = g_flops;
    auto attempt_ops = [&](double h0, bool* accepted) {
        DP5T<Flop> st{a, Flop(1e-9), Flop(1e-11), {}, false, Flop(1e-4), 0};
        rhs_t<Flop>(a, y, st.k1);
        st.k1_ok = true;
        Flop h = h0, yn[8];
        SegT<Flop> seg;
        g_flops = 0;
        *accepted = st.attempt(y, h, yn, seg);
        return g_flops - 6 * per_rhs;
    };
    bool acc_small = false, acc_big = true;
    const long long att_acc = attempt_ops(1e-3, &acc_small);
    const long long att_rej = attempt_ops(5.0, &acc_big);
    // per accepted step outside the attempt
    g_flops = 0;
    const InvT<Flop> iv = invariants_t<Flop>(a, y);
    Flop p1[3];
    to_cartesian_t<Flop>(a, y[YR], y[YH], y[YP], p1);
    const long long inv_ops = g_flops;
    const long long loop_ops = 3      // hmax = 0.05 max(1, r - r_h) + 0.01
                             + 2      // (Q - Q0) / (|Q0| + 1e-12): the drift of Carter's Q
                             + 3;     // cos(theta_0), cos(theta_1), their product
    (void)iv;
    out[0] = per_rhs;
    out[1] = acc_small ? att_acc : -1;
    out[2] = acc_big ? -1 : att_rej;
    out[3] = inv_ops + loop_ops;
    out[4] = 6 * per_rhs + out[1] + out[3];   // one accepted step
    out[5] = 6 * per_rhs + out[2];            // one rejected attempt
}


int oracle_render(const bhx_scene* s, const oracle_catalogue* cat, double* xyz, bhx_stats* stats,
                  int threads) {
    if (!s || !xyz || s->width <= 0 || s->height <= 0) return BHX_E_INVALID;
    const Scene sc(s, cat);
    const int S = std::max(1, s->sqrt_spp);
    const int W = s->width, H = s->height;
    std::vector<double> row_flux(H, 0.0);
    std::atomic<int> next_row{0};
    const int nt = n_threads(threads);
    std::vector<bhx_stats> per(nt);
    for (auto& p : per) std::memset(&p, 0, sizeof p);
    auto worker = [&](int tid) {
        bhx_stats& st = per[tid];
        for (;;) {
            const int py = next_row.fetch_add(1);
            if (py >= H) break;
            double flux = 0.0;
            for (int px = 0; px < W; ++px) {
                double acc[3] = {0.0, 0.0, 0.0};
                for (int k = 0; k < S * S; ++k) {
                    RayResult r;
                    trace_one(sc, (static_cast<int64_t>(py) * W + px) * S * S + k, r, st);
                    acc[0] += r.xyz[0]; acc[1] += r.xyz[1]; acc[2] += r.xyz[2];
                }
                const double w = 1.0 / (S * S);
                double* o = xyz + (static_cast<size_t>(py) * W + px) * 3;
                o[0] = acc[0] * w; o[1] = acc[1] * w; o[2] = acc[2] * w;
                flux += acc[1] * w;
            }
            row_flux[py] = flux;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(worker, t);
    for (auto& t : pool) t.join();
    if (stats) {
        bhx_stats tot;
        std::memset(&tot, 0, sizeof tot);
        for (auto& p : per) merge(tot, p);
        for (int py = 0; py < H; ++py) tot.total_flux += row_flux[py];
        merge(*stats, tot);
    }
    return BHX_OK;
}

int oracle_trace_list(const bhx_scene* s, const oracle_catalogue* cat, const int64_t* ids,
                      int64_t n, oracle_rays* out, int threads) {
    if (!s || !ids || !out || n < 0) return BHX_E_INVALID;
    const Scene sc(s, cat);
    std::atomic<int64_t> next{0};
    const int nt = n_threads(threads);
    auto worker = [&]() {
        bhx_stats st;
        std::memset(&st, 0, sizeof st);
        for (;;) {
            const int64_t i = next.fetch_add(1);
            if (i >= n) break;
            RayResult r;
            trace_one(sc, ids[i], r, st);
            if (out->cls) out->cls[i] = static_cast<uint8_t>(r.cls);
            if (out->y) for (int c = 0; c < 8; ++c) out->y[c * n + i] = r.y[c];
            if (out->jet) out->jet[i] = r.jet;
            if (out->steps) out->steps[i] = r.steps;
            if (out->attempts) out->attempts[i] = r.attempts;
            if (out->norm_err) out->norm_err[i] = r.norm;
            if (out->carter) out->carter[i] = r.carter;
            if (out->g) out->g[i] = r.g;
            if (out->dir) for (int c = 0; c < 3; ++c) out->dir[c * n + i] = r.dir[c];
            if (out->ray_xyz) for (int c = 0; c < 3; ++c) out->ray_xyz[c * n + i] = r.xyz[c];
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return BHX_OK;
}

int oracle_trace_uv(const bhx_scene* s, const oracle_catalogue* cat, const double* u, const double* v,
                    int64_t n, oracle_rays* out, int threads) {
    if (!s || !u || !v || !out || n < 0) return BHX_E_INVALID;
    const Scene sc(s, cat);
    std::atomic<int64_t> next{0};
    const int nt = n_threads(threads);
    auto worker = [&]() {
        bhx_stats st;
        std::memset(&st, 0, sizeof st);
        for (;;) {
            const int64_t i = next.fetch_add(1);
            if (i >= n) break;
            double dir[3], y0[8];
            sc.ray_dir(u[i], v[i], dir);
            photon(sc.a, sc.c.cam_r, sc.c.cam_theta, sc.c.cam_phi, sc.tet, dir, y0);
            RayResult r;
            trace_state(sc, y0, r);
            shade_ray(sc, r, st);
            if (out->cls) out->cls[i] = static_cast<uint8_t>(r.cls);
            if (out->y) for (int c = 0; c < 8; ++c) out->y[c * n + i] = r.y[c];
            if (out->jet) out->jet[i] = r.jet;
            if (out->steps) out->steps[i] = r.steps;
            if (out->attempts) out->attempts[i] = r.attempts;
            if (out->norm_err) out->norm_err[i] = r.norm;
            if (out->carter) out->carter[i] = r.carter;
            if (out->g) out->g[i] = r.g;
            if (out->dir) for (int c = 0; c < 3; ++c) out->dir[c * n + i] = r.dir[c];
            if (out->ray_xyz) for (int c = 0; c < 3; ++c) out->ray_xyz[c * n + i] = r.xyz[c];
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    return BHX_OK;
}

// Image-plane coordinates (u, v) of each ray (render.cpp:445-448).
int oracle_ray_uv(const bhx_scene* s, const int64_t* ids, int64_t n, double* u, double* v) {
    const int S = std::max(1, s->sqrt_spp);
    const double inv_S = 1.0 / S;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t pix = ids[i] / (S * S);
        const int k = static_cast<int>(ids[i] % (S * S));
        const int px = static_cast<int>(pix % s->width), py = static_cast<int>(pix / s->width);
        SplitMix rng(s->seed ^ (static_cast<uint64_t>(py) << 32) ^
                     (static_cast<uint64_t>(px) * 0x9E3779B97F4A7C15ull));
        double jx = 0.0, jy = 0.0;
        for (int q = 0; q <= k; ++q) {
            jx = ((q % S) + rng.u01()) * inv_S;
            jy = ((q / S) + rng.u01()) * inv_S;
        }
        u[i] = (2.0 * (px + jx) / s->width - 1.0);
        v[i] = (1.0 - 2.0 * (py + jy) / s->height);
    }
    return BHX_OK;
}

int oracle_trace_state(const bhx_scene* s, const double* y0, double* y_out, double* jet,
                       int32_t* steps, double* norm_err, double* carter) {
    bhx_scene c = *s;
    c.stars_enabled = 0;
    const Scene sc(&c, nullptr);
    RayResult r;
    trace_state(sc, y0, r);
    if (y_out) std::memcpy(y_out, r.y, sizeof r.y);
    if (jet) *jet = r.jet;
    if (steps) *steps = r.steps;
    if (norm_err) *norm_err = r.norm;
    if (carter) *carter = r.carter;
    return r.cls;
}

void oracle_geodesic_rhs(double a, const double* y, double* dy) { rhs(a, y, dy); }

void oracle_make_photon(double a, double r, double th, double phi, const double* dir, double* y) {
    double e[4][4];
    zamo(point(a, r, th), e);
    photon(a, r, th, phi, e, dir, y);
}

void oracle_invariants(double a, const double* y, double* E, double* Lz, double* Q, double* norm,
                       double* norm_rel) {
    const Inv v = invariants(a, y);
    if (E) *E = v.E;
    if (Lz) *Lz = v.Lz;
    if (Q) *Q = v.Q;
    if (norm) *norm = v.norm;
    if (norm_rel) *norm_rel = v.norm_rel;
}

int oracle_integrate(double a, double* y, double h0, double rtol, double atol, int max_steps,
                     double r_min, double r_max, int64_t* nsteps) {
    DP5 st{a, rtol, atol, {}, false, 1e-4, 0};
    st.reset();
    double h = h0, yn[8];
    Seg seg;
    for (int i = 0; i < max_steps; ++i) {
        int tries = 0;
        while (!st.attempt(y, h, yn, seg))
            if (++tries > 60) return 1;
        if (nsteps) ++*nsteps;
        if (yn[YR] < r_min || yn[YR] > r_max) { std::memcpy(y, yn, sizeof yn); return 0; }
        std::memcpy(y, yn, sizeof yn);
    }
    return 2;
}

double oracle_isco(double a, int sense) { return isco(a, sense); }
double oracle_page_thorne_W(double a, double r, double r_isco) { return pt_W(a, r, r_isco); }
double oracle_disc_temperature(const bhx_scene* s, double r) {
    const Disc d(s->M_kg, s->spin, s->disc_r_out, s->eddington, s->disc_sense);
    return d.temperature(r);
}
// Bolometric intensity of the disc hits among n rays (class 1 with g > 0 and
// T > 0), else 0: spec::blackbody_bolometric (spectrum.cpp:24-26) of the
// observed temperature g T, sigma T^4 / pi in the reference's operation
// order; the loss input of the differentiable disc image.
void oracle_disc_intensity(const bhx_scene* s, int64_t n, const uint8_t* cls, const double* r, const double* g,
                           double* out) {
    const Disc d(s->M_kg, s->spin, s->disc_r_out, s->eddington, s->disc_sense);
    for (int64_t i = 0; i < n; ++i) {
        out[i] = 0.0;
        if (cls[i] != 1 || !(g[i] > 0.0)) continue;
        const double T = d.temperature(r[i]);
        if (!(T > 0.0)) continue;
        const double gT = g[i] * T;
        out[i] = kSigmaSB * gT * gT * gT * gT / kPi;
    }
}
int oracle_uniform_sky(int32_t count, uint64_t seed, double* x, double* y, double* z, double* T, double* irr) {
    if (count < 0) return -1;
    Sky sk;
    uniform_sky(count, seed, sk);
    std::copy(sk.x.begin(), sk.x.end(), x);
    std::copy(sk.y.begin(), sk.y.end(), y);
    std::copy(sk.z.begin(), sk.z.end(), z);
    std::copy(sk.T.begin(), sk.T.end(), T);
    std::copy(sk.irr.begin(), sk.irr.end(), irr);
    return 0;
}

// ---- conical jet pieces (this package's own jet model), for the checks ----
// Four-velocity of the jet's radial outflow at (a, r, theta) for a Lorentz
// factor gamma relative to the zero-angular-momentum observer.
void oracle_cone_velocity(double a, double r, double th, double gamma, double* u) {
    JetModel j{};
    j.gamma_max = gamma;
    j.cone_velocity(point(a, r, th), u);
}
// Emissivity profile of the scene's conical jet at (r, cos theta), n points.
void oracle_cone_shape(const bhx_scene* s, int64_t n, const double* r, const double* cs, double* out) {
    JetModel j{};
    j.model = 1;
    j.cone_angle = s->jet_cone_angle;
    j.cone_r0 = s->jet_cone_r0;
    j.cone_r1 = s->jet_cone_r1;
    for (int64_t i = 0; i < n; ++i) out[i] = j.cone_shape(r[i], cs[i]);
}
// The sampling gate of the conical jet for n steps whose dense-output theta
// coefficients are c[k * n + i], k = 0..4 (c1..c5 of the Hairer interpolant).
void oracle_cone_gate(const bhx_scene* s, int64_t n, const double* c, uint8_t* out) {
    JetModel j{};
    j.model = 1;
    j.cone_angle = s->jet_cone_angle;
    for (int64_t i = 0; i < n; ++i)
        out[i] = j.cone_gate(c[i], c[n + i], c[2 * n + i], c[3 * n + i], c[4 * n + i]) ? 1 : 0;
}
// The redshift factor g = 1 / (p . u) of a photon state for an emitter of
// four-velocity u (render.cpp:31-34), and the jet's boost g^(2 + alpha).
double oracle_redshift(const double* y, const double* u) { return redshift(y, u); }
double oracle_jet_boost(double alpha, double g) {
    JetModel j{};
    j.alpha = alpha;
    return j.boost(g);
}

void oracle_blackbody_xyz(double T, double* o) { bb_xyz(T, o); }
void oracle_blackbody_xyz_exact(double T, double* o) { bb_exact(T, o); }
void oracle_powerlaw_xyz(double alpha, double* o) { powerlaw(alpha, o); }

void oracle_star_radiance(const bhx_scene* s, const oracle_catalogue* cat, const double* d,
                          double sigma, double* o) {
    Sky sk;
    if (cat && cat->n > 0) {
        sk.x.assign(cat->x, cat->x + cat->n);
        sk.y.assign(cat->y, cat->y + cat->n);
        sk.z.assign(cat->z, cat->z + cat->n);
        sk.T.assign(cat->temperature, cat->temperature + cat->n);
        sk.irr.assign(cat->irradiance, cat->irradiance + cat->n);
        sk.bin_all();
    } else {
        generate_sky(s->star_count, s->seed, sk);
    }
    sky_radiance(sk, d, sigma, o);
}

int oracle_generate_catalogue(int32_t count, uint64_t seed, double* x, double* y, double* z,
                              double* T, double* irr) {
    Sky sk;
    generate_sky(count, seed, sk);
    for (int i = 0; i < count; ++i) {
        x[i] = sk.x[i]; y[i] = sk.y[i]; z[i] = sk.z[i]; T[i] = sk.T[i]; irr[i] = sk.irr[i];
    }
    return BHX_OK;
}

// img::add_glare (image.cpp:47-106): the image plus a narrow (sigma = R) and
// a wide (3R) separable Gaussian of itself, weights 1 - s, 0.7 s, 0.3 s.
// Each blur: taps -r..r, r = max(1, ceil(3 sigma)), normalised weights,
// clamp to the edge; horizontal pass then vertical pass, sums in tap order.
void oracle_add_glare(double* xyz, int32_t w, int32_t h, double strength, double radius_px) {
    if (strength <= 0.0 || radius_px <= 0.0) return;
    const size_t n = static_cast<size_t>(w) * h;
    auto gauss = [&](double sigma, std::vector<double>& out) {
        const int r = std::max(1, static_cast<int>(std::ceil(3.0 * sigma)));
        std::vector<double> k(2 * r + 1);
        double sum = 0.0;
        for (int i = -r; i <= r; ++i) {
            k[i + r] = std::exp(-0.5 * (i * i) / (sigma * sigma));
            sum += k[i + r];
        }
        for (double& v : k) v /= sum;
        std::vector<double> rows(3 * n);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) {
                    double a = 0.0;
                    for (int i = -r; i <= r; ++i) {
                        const int xs = std::min(std::max(x + i, 0), w - 1);
                        a += k[i + r] * xyz[3 * (static_cast<size_t>(y) * w + xs) + c];
                    }
                    rows[3 * (static_cast<size_t>(y) * w + x) + c] = a;
                }
        out.assign(3 * n, 0.0);
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) {
                    double a = 0.0;
                    for (int i = -r; i <= r; ++i) {
                        const int ys = std::min(std::max(y + i, 0), h - 1);
                        a += k[i + r] * rows[3 * (static_cast<size_t>(ys) * w + x) + c];
                    }
                    out[3 * (static_cast<size_t>(y) * w + x) + c] = a;
                }
    };
    std::vector<double> nearb, farb;
    gauss(radius_px, nearb);
    gauss(3.0 * radius_px, farb);
    const double core = 1.0 - strength, wa = 0.7 * strength, wb = 0.3 * strength;
    for (size_t i = 0; i < 3 * n; ++i) xyz[i] = xyz[i] * core + nearb[i] * wa + farb[i] * wb;
}

// img::auto_exposure (image.cpp:22-45)
double oracle_auto_exposure(const double* xyz, int64_t n, double key) {
    std::vector<double> lum;
    lum.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        const double Y = xyz[i * 3 + 1];
        if (Y > 0.0 && std::isfinite(Y)) lum.push_back(Y);
    }
    if (lum.empty()) return 1.0;
    auto pct = [&](double q) {
        size_t k = static_cast<size_t>(q * (lum.size() - 1));
        k = std::min(k, lum.size() - 1);
        std::nth_element(lum.begin(), lum.begin() + k, lum.end());
        return lum[k];
    };
    const double hi = pct(0.999);
    const double peak = pct(0.9999);
    double e = 1.0;
    if (hi > 0.0) e = key / hi;
    if (peak > 0.0) e = std::min(e, 3.0 * key / peak);
    return e;
}

// img::develop (image.cpp:109-191) with spec::xyz_to_linear_srgb / gamut_clamp
// (spectrum.cpp:137-152).
void oracle_develop(const double* xyz, int64_t n, double exposure, int tm, double log_decades,
                    uint8_t* rgb) {
    auto aces = [](double x) {
        const double A = 2.51, B = 0.03, C = 2.43, D = 0.59, E = 0.14;
        return std::clamp((x * (A * x + B)) / (x * (C * x + D) + E), 0.0, 1.0);
    };
    auto oetf = [](double v) {
        v = std::clamp(v, 0.0, 1.0);
        return (v <= 0.0031308) ? 12.92 * v : 1.055 * std::pow(v, 1.0 / 2.4) - 0.055;
    };
    const double log_k = std::pow(10.0, std::clamp(log_decades, 1.0, 20.0));
    for (int64_t i = 0; i < n; ++i) {
        double c[3] = {xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]};
        if (!std::isfinite(c[0]) || !std::isfinite(c[1]) || !std::isfinite(c[2]))
            c[0] = c[1] = c[2] = 0.0;
        c[0] *= exposure; c[1] *= exposure; c[2] *= exposure;
        double v[3];
        v[0] = 3.2404542 * c[0] - 1.5371385 * c[1] - 0.4985314 * c[2];
        v[1] = -0.9692660 * c[0] + 1.8760108 * c[1] + 0.0415560 * c[2];
        v[2] = 0.0556434 * c[0] - 0.2040259 * c[1] + 1.0572252 * c[2];
        const double m = std::min({v[0], v[1], v[2]});
        if (!(m >= 0.0)) { v[0] -= m; v[1] -= m; v[2] -= m; }
        v[0] = std::max(v[0], 0.0); v[1] = std::max(v[1], 0.0); v[2] = std::max(v[2], 0.0);
        switch (tm) {
            case BHX_TM_LINEAR:
                for (double& x : v) x = std::clamp(x, 0.0, 1.0);
                break;
            case BHX_TM_REINHARD:
                for (double& x : v) x = x / (1.0 + x);
                break;
            case BHX_TM_ACES:
                for (double& x : v) x = aces(x);
                break;
            case BHX_TM_HDR: {
                const double cap = 0.12, w = 0.5;
                const double denom = std::log10(1.0 + log_k);
                for (double& x : v) {
                    const double f = aces(x);
                    const double l =
                        std::clamp(std::log10(1.0 + log_k * std::max(x, 0.0)) / denom, 0.0, 1.0);
                    const double lift = cap * std::tanh(w * l / cap);
                    x = std::clamp(f + lift * (1.0 - f), 0.0, 1.0);
                }
                break;
            }
            case BHX_TM_LOG: {
                const double denom = std::log10(1.0 + log_k);
                for (double& x : v)
                    x = std::clamp(std::log10(1.0 + log_k * std::max(x, 0.0)) / denom, 0.0, 1.0);
                break;
            }
        }
        for (int ch = 0; ch < 3; ++ch)
            rgb[i * 3 + ch] = static_cast<uint8_t>(std::lround(255.0 * oetf(v[ch])));
    }
}

}  // extern "C"
