// host_setup.cpp - host-side scene preparation (see host_setup.h).
//
// Scalars that the reference evaluates once per frame on the CPU are computed
// here with the host libm, in the reference's own expression order, so the
// kernels start from the same constants the reference would use.
#include "host_setup.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <thread>

namespace bhx {
namespace host {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kG = 6.67430e-11, kC = 2.99792458e8, kH = 6.62607015e-34, kKB = 1.380649e-23;
constexpr double kSigmaT = 6.6524587321e-29, kMp = 1.67262192369e-27;
constexpr int kTabN = 4096;
constexpr double kLgMin = 1.6989700043360187, kLgMax = 9.4771212547196626;

// Planck B_lambda (spectrum.cpp:13-22) and the Wyman-Sloan-Shirley CMF fit
// (spectrum.cpp:36-53).
double planck(double lam, double T) {
    if (T <= 0.0 || lam <= 0.0) return 0.0;
    const double l5 = lam * lam * lam * lam * lam;
    const double x = kH * kC / (lam * kKB * T);
    if (x > 700.0) return 0.0;
    return (2.0 * kH * kC * kC) / (l5 * std::expm1(x));
}
double lobe(double x, double mu, double s1, double s2) {
    const double t = (x - mu) * ((x < mu) ? 1.0 / s1 : 1.0 / s2);
    return std::exp(-0.5 * t * t);
}
void cmf(double nm, double o[3]) {
    o[0] = 1.056 * lobe(nm, 599.8, 37.9, 31.0) + 0.362 * lobe(nm, 442.0, 16.0, 26.7) -
           0.065 * lobe(nm, 501.1, 20.4, 26.2);
    o[1] = 0.821 * lobe(nm, 568.8, 46.9, 40.5) + 0.286 * lobe(nm, 530.9, 16.3, 31.1);
    o[2] = 1.217 * lobe(nm, 437.0, 11.8, 36.0) + 0.681 * lobe(nm, 459.0, 26.0, 13.8);
}
void bb_exact(double T, double o[3]) {
    o[0] = o[1] = o[2] = 0.0;
    if (T <= 0.0) return;
    for (double nm = 360.0; nm <= 830.0; nm += 1.0) {
        const double w = planck(nm * 1e-9, T) * 1.0 * 1e-9;
        double c[3];
        cmf(nm, c);
        o[0] += w * c[0];
        o[1] += w * c[1];
        o[2] += w * c[2];
    }
}

struct SplitMix {
    uint64_t s;
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

double bv_to_T(double bv) {  // Ballesteros (2012), spectrum.cpp:158-162
    const double x = std::clamp(bv, -0.4, 2.0);
    return 4600.0 * (1.0 / (0.92 * x + 1.7) + 1.0 / (0.92 * x + 0.62));
}

double circ_energy(double a, double r, int sg) {  // kerr.h:232-237
    const double sr = std::sqrt(r);
    return (r * r - 2.0 * r + sg * a * sr) / (r * std::sqrt(r * r - 3.0 * r + 2.0 * sg * a * sr));
}

}  // namespace

double horizon_outer(double a) { return 1.0 + std::sqrt(std::max(0.0, 1.0 - a * a)); }
double horizon_inner(double a) { return 1.0 - std::sqrt(std::max(0.0, 1.0 - a * a)); }

double isco_radius(double a, int sense) {  // Bardeen, Press & Teukolsky (1972) eq. 2.21
    const double a2 = a * a;
    const double Z1 = 1.0 + std::cbrt(1.0 - a2) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    const double Z2 = std::sqrt(3.0 * a2 + Z1 * Z1);
    const double s = std::sqrt(std::max(0.0, (3.0 - Z1) * (3.0 + Z1 + 2.0 * Z2)));
    return 3.0 + Z2 - static_cast<double>(sense) * s;
}

const std::vector<double>& bb_table() {
    static std::vector<double> tab;
    static std::once_flag once;
    std::call_once(once, [] {
        tab.resize(3 * kTabN);
        const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t)
            pool.emplace_back([t, nt] {
                for (int i = static_cast<int>(t); i < kTabN; i += static_cast<int>(nt)) {
                    const double f = static_cast<double>(i) / (kTabN - 1);
                    const double T = std::pow(10.0, kLgMin + f * (kLgMax - kLgMin));
                    double v[3];
                    bb_exact(T, v);
                    for (int c = 0; c < 3; ++c) tab[c * kTabN + i] = std::log(std::max(v[c], 1e-300));
                }
            });
        for (auto& th : pool) th.join();
    });
    return tab;
}

void blackbody_xyz(double T, double o[3]) {
    const std::vector<double>& t = bb_table();
    const double lg = std::log10(std::clamp(T, 50.0, 3.0e9));
    const double f = (lg - kLgMin) / (kLgMax - kLgMin) * (kTabN - 1);
    const int i = std::clamp(static_cast<int>(f), 0, kTabN - 2);
    const double u = f - i;
    for (int c = 0; c < 3; ++c)
        o[c] = std::exp(t[c * kTabN + i] * (1.0 - u) + t[c * kTabN + i + 1] * u);
}

void powerlaw_xyz(double alpha, double o[3]) {
    const double lambda0 = 550e-9;
    o[0] = o[1] = o[2] = 0.0;
    for (double nm = 360.0; nm <= 830.0; nm += 1.0) {
        const double lam = nm * 1e-9;
        const double Il = kC * std::pow(lam / lambda0, alpha) / (lam * lam);
        const double w = Il * 1.0 * 1e-9;
        double c[3];
        cmf(nm, c);
        o[0] += w * c[0];
        o[1] += w * c[1];
        o[2] += w * c[2];
    }
}

void generate_catalogue(int count, uint64_t seed, Catalogue& cat) {
    SplitMix rng{seed};
    const double ca = 0.8660254037844387, sa = 0.5;  // galactic plane tilted by 30 degrees
    cat.x.resize(count); cat.y.resize(count); cat.z.resize(count);
    cat.T.resize(count); cat.irr.resize(count);
    for (int i = 0; i < count; ++i) {
        double d[3];
        if (rng.u01() < 0.45) {  // 45% drawn towards the galactic plane
            const double b = rng.gauss() * 0.12;
            const double l = rng.u(0.0, 2.0 * kPi);
            const double cb = std::cos(b);
            const double g0 = cb * std::cos(l), g1 = cb * std::sin(l), g2 = std::sin(b);
            d[0] = g0;
            d[1] = ca * g1 - sa * g2;
            d[2] = sa * g1 + ca * g2;
        } else {  // isotropic
            const double z = rng.u(-1.0, 1.0);
            const double ph = rng.u(0.0, 2.0 * kPi);
            const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
            d[0] = rr * std::cos(ph);
            d[1] = rr * std::sin(ph);
            d[2] = z;
        }
        cat.x[i] = d[0]; cat.y[i] = d[1]; cat.z[i] = d[2];
        // inverse-CDF magnitude of N(<m) ~ 10^(0.45 m) on [-1.5, 12]
        const double k = 0.45, m_lo = -1.5, m_hi = 12.0;
        const double u = rng.u01();
        const double mag = m_lo + std::log10(u * (std::pow(10.0, k * (m_hi - m_lo)) - 1.0) + 1.0) / k;
        double bv = 0.55 + 0.42 * rng.gauss();
        if (rng.u01() < 0.10) bv = rng.u(1.2, 1.9);  // red-giant tail
        cat.T[i] = bv_to_T(std::clamp(bv, -0.35, 1.95));
        const double F0 = 3.64e-9 * 0.089 * 1e6 * 1e-6;  // V = 0 irradiance, W m^-2
        cat.irr[i] = F0 * std::pow(10.0, -0.4 * mag);
    }
}

void uniform_sky(int count, uint64_t seed, Catalogue& cat) {
    SplitMix rng{seed};
    const double k = 0.45, m_lo = -1.5, m_hi = 8.0;
    const double span = std::pow(10.0, k * (m_hi - m_lo)) - 1.0;
    const double F0 = 3.64e-9 * 0.089;   // V = 0: 3.64e-9 W m^-2 um^-1 over 0.089 um
    cat.x.resize(count); cat.y.resize(count); cat.z.resize(count);
    cat.T.resize(count); cat.irr.resize(count);
    for (int i = 0; i < count; ++i) {
        // uniform on the sphere: z uniform on [-1, 1] (Archimedes), phi on [0, 2 pi)
        const double z = 2.0 * rng.u01() - 1.0;
        const double ph = 2.0 * kPi * rng.u01();
        const double rho = std::sqrt(std::max(0.0, (1.0 - z) * (1.0 + z)));
        cat.x[i] = rho * std::cos(ph);
        cat.y[i] = rho * std::sin(ph);
        cat.z[i] = z;
        // magnitude by inverse CDF of N(<m) ~ 10^(k m) on [m_lo, m_hi]
        const double m = m_lo + std::log10(1.0 + span * rng.u01()) / k;
        // B-V ~ N(0.6, 0.4) by Box-Muller (one of the pair)
        const double radius = std::sqrt(-2.0 * std::log(1.0 - rng.u01()));
        const double bv = 0.6 + 0.4 * radius * std::cos(2.0 * kPi * rng.u01());
        cat.T[i] = bv_to_T(bv);
        cat.irr[i] = F0 * std::pow(10.0, -0.4 * m);
    }
}

void bin_catalogue(const Catalogue& cat, std::vector<uint32_t>& cell_start, std::vector<double>& rows) {
    const int n_lat = 360, n_lon = 720, n_cells = n_lat * n_lon;
    const size_t n = cat.x.size();
    std::vector<uint32_t> cell(n);
    cell_start.assign(n_cells + 1, 0u);
    for (size_t i = 0; i < n; ++i) {
        const double th = std::acos(std::clamp(cat.z[i], -1.0, 1.0));
        const double ph = std::atan2(cat.y[i], cat.x[i]);
        int ti = static_cast<int>(th / kPi * n_lat);
        int pi = static_cast<int>((ph + kPi) / (2.0 * kPi) * n_lon);
        ti = std::clamp(ti, 0, n_lat - 1);
        pi = ((pi % n_lon) + n_lon) % n_lon;
        cell[i] = static_cast<uint32_t>(ti * n_lon + pi);
        ++cell_start[cell[i] + 1];
    }
    for (int c = 0; c < n_cells; ++c) cell_start[c + 1] += cell_start[c];
    std::vector<uint32_t> fill(cell_start.begin(), cell_start.end() - 1);
    rows.assign(6 * n, 0.0);
    for (size_t i = 0; i < n; ++i) {  // stable: ascending star index within a cell
        const uint32_t slot = fill[cell[i]]++;
        double c[3];
        blackbody_xyz(cat.T[i], c);
        const double inv = (c[1] > 0.0) ? 1.0 / c[1] : 0.0;
        double* r = &rows[6 * static_cast<size_t>(slot)];
        r[0] = cat.x[i]; r[1] = cat.y[i]; r[2] = cat.z[i]; r[3] = cat.irr[i];
        r[4] = c[0] * inv; r[5] = c[2] * inv;
    }
}

RaygenArgs raygen_args(const bhx_scene& s) {
    RaygenArgs P{};
    const double a = s.spin;
    P.a = a;
    P.a2 = a * a;
    // camera point: Geom, ZAMO tetrad, covariant metric (kerr.h:44-77, 270-293)
    const double r = s.cam_r, th = s.cam_theta;
    double sn = std::sin(th);
    if (std::fabs(sn) < 1e-9) sn = (sn < 0.0) ? -1e-9 : 1e-9;
    const double cs = std::cos(th);
    const double sn2 = sn * sn, cs2 = cs * cs;
    const double S = r * r + P.a2 * cs2;
    const double D = r * r - 2.0 * r + P.a2;
    const double A = (r * r + P.a2) * (r * r + P.a2) - P.a2 * D * sn2;
    const double alpha = std::sqrt(S * D / A);
    const double omega = 2.0 * a * r / A;
    P.cam_r = r;
    P.cam_th = th;
    P.cam_ph = s.cam_phi;
    P.e00 = 1.0 / alpha;
    P.e03 = omega / alpha;
    P.e11 = std::sqrt(D / S);
    P.e22 = 1.0 / std::sqrt(S);
    P.e33 = std::sqrt(S / A) / sn;
    P.gl_tt = -(1.0 - 2.0 * r / S);
    P.gl_tp = -2.0 * a * r * sn2 / S;
    P.gl_rr = S / D;
    P.gl_hh = S;
    P.gl_pp = A * sn2 / S;
    double snq = std::sin(th);  // carter_constant clamps without the sign (kerr.h:344-345)
    if (std::fabs(snq) < 1e-9) snq = 1e-9;
    P.cam_cs2 = std::cos(th) * std::cos(th);
    P.cam_sn2 = snq * snq;
    P.tanf = std::tan(0.5 * s.cam_fov);
    P.aspect = static_cast<double>(s.width) / s.height;
    const int Sspp = std::max(1, s.sqrt_spp);
    P.inv_S = 1.0 / Sspp;
    P.has_roll = s.cam_roll != 0.0;
    P.c_roll = std::cos(s.cam_roll); P.s_roll = std::sin(s.cam_roll);
    P.c_pitch = std::cos(s.cam_pitch); P.s_pitch = std::sin(s.cam_pitch);
    P.c_yaw = std::cos(s.cam_yaw); P.s_yaw = std::sin(s.cam_yaw);
    P.W = s.width;
    P.H = s.height;
    P.S = Sspp;
    P.seed = s.seed;
    P.n_rays = s.width * s.height * Sspp * Sspp;
    return P;
}

EventParams event_params(const bhx_scene& s) {
    EventParams E{};
    E.z_base = s.jet_z_base; E.z_top = s.jet_z_top; E.R_base = s.jet_R_base; E.k = s.jet_k;
    E.gamma_max = s.jet_gamma_max; E.z_sat = s.jet_z_sat; E.alpha = s.jet_alpha;
    E.u_peak = s.jet_u_peak; E.u_wid = s.jet_u_wid; E.spine = s.jet_spine;
    E.spine_wid = s.jet_spine_wid; E.zeta = s.jet_zeta; E.Omega_F = s.jet_Omega_F;
    E.j_nu0 = s.jet_j_nu0; E.r_bound = s.jet_r_bound;
    // Jet::gamma_at(z_base) (jet.h:159-162)
    const double t = std::min(1.0, std::pow(std::max(s.jet_z_base, 1e-6) / s.jet_z_sat, s.jet_k));
    E.gamma_zb = 1.0 + (s.jet_gamma_max - 1.0) * t;
    for (int i = 0; i < 3; ++i) E.sun_c[i] = s.sun_centre[i];
    E.sun_R = s.sun_radius;
    E.model = s.jet_model;
    E.cone_angle = s.jet_cone_angle;
    E.cone_r0 = s.jet_cone_r0;
    E.cone_r1 = s.jet_cone_r1;
    E.cone_gamma = std::max(1.0, s.jet_gamma_max);
    E.cone_beta = std::sqrt(std::max(0.0, 1.0 - 1.0 / (E.cone_gamma * E.cone_gamma)));
    E.cone_cos = std::cos(s.jet_cone_angle);
    E.cone_itaper = 1.0 / (std::cos(0.8 * s.jet_cone_angle) - E.cone_cos);
    return E;
}

void tile_queue(int W, int H, int S, int32_t q_base, std::vector<int32_t>& order,
                std::vector<int32_t>& start, bool axis_band, double roll) {
    const int tw = (W + 7) / 8, th = (H + 7) / 8;
    const int n = tw * th;
    order.resize(n);
    std::vector<double> key(n);
    const double cr = std::cos(roll), sr = std::sin(roll);
    for (int i = 0; i < n; ++i) {
        order[i] = i;
        const double cx = (i % tw) * 8 + 4 - 0.5 * W, cy = (i / tw) * 8 + 4 - 0.5 * H;
        key[i] = cx * cx + cy * cy;
        if (axis_band) {   // the camera's x (after roll) vanishes on the axis' image
            const double perp = cr * cx + sr * cy;
            key[i] = std::min(key[i], perp * perp);
        }
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    start.assign(n + 1, q_base);
    for (int t = 0; t < n; ++t) {
        const int tx = order[t] % tw, ty = order[t] / tw;
        start[t + 1] = start[t] + std::min(8, W - tx * 8) * std::min(8, H - ty * 8) * S * S;
    }
}

TraceArgs trace_args(const bhx_scene& s) {
    TraceArgs P{};
    const double a = s.spin;
    P.a = a;
    P.a2 = a * a;
    P.r_h = horizon_outer(a);
    P.r_stop = P.r_h * 1.0 + 1e-4 + (P.r_h - horizon_inner(a)) * 1e-3;  // render.cpp:72-73
    P.rtol = s.rtol;
    P.atol = s.atol;
    P.r_escape = s.r_escape;
    P.h0 = 0.02 * std::max(1.0, s.cam_r);                              // render.cpp:86
    P.max_steps = s.max_steps;
    P.has_disc = s.disc_enabled != 0;
    if (P.has_disc) {  // NovikovThorneDisc ctor (disc.h:73-82)
        const double ac = std::clamp(s.spin, -0.9999, 0.9999);
        P.r_in = isco_radius(ac, s.disc_sense);
        P.r_out = std::max(s.disc_r_out, isco_radius(ac, s.disc_sense) * 1.02);
    }
    P.has_ns = s.ns_enabled != 0;
    P.ns_R = s.ns_R;
    P.has_jet = s.jet_enabled != 0;
    P.has_sun = s.sun_enabled != 0;
    const int Sspp = std::max(1, s.sqrt_spp);
    P.n_rays = s.width * s.height * Sspp * Sspp;
    return P;
}

ShadeArgs shade_args(const bhx_scene& s) {
    ShadeArgs P{};
    P.a = s.spin;
    P.a2 = s.spin * s.spin;
    P.psf_sigma = 0.55 * (s.cam_fov / s.width);  // render.cpp:414-415
    P.observer_time = s.observer_time;
    P.W = s.width;
    P.H = s.height;
    P.S = std::max(1, s.sqrt_spp);
    P.n_pix = s.width * s.height;
    P.n_rays = static_cast<int64_t>(P.n_pix) * P.S * P.S;
    P.has_disc = s.disc_enabled != 0;
    if (P.has_disc) {
        const double a = std::clamp(s.spin, -0.9999, 0.9999);
        const int sg = s.disc_sense;
        P.disc_a = a;
        P.disc_sg = sg;
        P.r_in = isco_radius(a, sg);
        P.r_out = std::max(s.disc_r_out, isco_radius(a, sg) * 1.02);
        const double eta = 1.0 - circ_energy(a, isco_radius(a, sg), sg);             // disc.h:78
        const double L_edd = 4.0 * kPi * kG * s.M_kg * kMp * kC / kSigmaT;          // constants.h:56-58
        const double mdot = s.eddington * (L_edd / (eta * kC * kC));                 // constants.h:62-64
        const double rg = kG * s.M_kg / (kC * kC);
        P.flux_scale = mdot * kC * kC / (rg * rg);
        // Page-Thorne roots of x^3 - 3x + 2a = 0 and their coefficients (disc.h:46-64)
        const double ac = std::acos(std::clamp(a, -1.0, 1.0));
        const double x[3] = {2.0 * std::cos(ac / 3.0 - kPi / 3.0), 2.0 * std::cos(ac / 3.0 + kPi / 3.0),
                             -2.0 * std::cos(ac / 3.0)};
        P.pt_x0 = std::sqrt(P.r_in);
        const int jj[3] = {1, 0, 0}, kk[3] = {2, 2, 1};  // term(x1,x2,x3), term(x2,x1,x3), term(x3,x1,x2)
        for (int i = 0; i < 3; ++i) {
            const double xi = x[i], xj = x[jj[i]], xk = x[kk[i]];
            P.pt_x[i] = xi;
            P.pt_c[i] = (std::fabs(xi) < 1e-13) ? 0.0 : 3.0 * (xi - a) * (xi - a) / (xi * (xi - xj) * (xi - xk));
        }
    }
    P.hs_enabled = s.hs_enabled;
    P.hs_r = s.hs_r; P.hs_phi0 = s.hs_phi0; P.hs_sigma = s.hs_sigma;
    P.hs_contrast = s.hs_contrast; P.hs_omega = s.hs_omega;
    P.ns_enabled = s.ns_enabled; P.ns_caps = s.ns_caps; P.ns_two_caps = s.ns_two_caps;
    P.ns_R = s.ns_R; P.ns_T_eff = s.ns_T_eff; P.ns_Omega = s.ns_Omega; P.ns_cap_T = s.ns_cap_T;
    P.ns_cap_radius = s.ns_cap_radius; P.ns_cap_phase0 = s.ns_cap_phase0;
    P.ns_sin_tilt = std::sin(s.ns_cap_tilt);
    P.ns_cos_tilt = std::cos(s.ns_cap_tilt);
    P.stars_enabled = s.stars_enabled;
    P.n_cells = 360 * 720;
    {
        const double F0 = 3.24e-10, per_sr = 4.2545170296152206e10;
        P.mw_L = F0 * std::pow(10.0, -0.4 * 21.5) * per_sr;
        double c[3];
        blackbody_xyz(4800.0, c);
        const double inv = (c[1] > 0.0) ? 1.0 / c[1] : 0.0;
        P.mw_cx = c[0] * inv;
        P.mw_cz = c[2] * inv;
    }
    if (s.jet_enabled) powerlaw_xyz(s.jet_alpha, P.jet_rgb);
    P.has_sun = s.sun_enabled;
    for (int i = 0; i < 3; ++i) P.sun_c[i] = s.sun_centre[i];
    P.sun_R = s.sun_radius;
    P.sun_T_eff = s.sun_T_eff;
    P.sun_self_z = kG * 1.98847e30 / (6.957e8 * kC * kC);   // G M_sun / (R_sun c^2), scene.h:239-240
    return P;
}

}  // namespace host
}  // namespace bhx
