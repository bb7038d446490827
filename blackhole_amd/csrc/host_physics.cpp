// host_physics.cpp - host solvers that prepare a scene before tracing.
//
//   bhx_setup_jet          emissivity of the reference's jet model (the jet of
//                          src/jet.h; its set-up src/main.cpp:399-423 and
//                          Jet::configure src/jet.h:261-311): the radiated
//                          fraction of the Blandford-Znajek power spread over
//                          the jets' emissivity-weighted volume
//   bhx_setup_conical_jet  the same for this package's conical jet
//   bhx_polytrope_star     stable-branch star of a given mass for P = K rho^Gamma
//                          (what ns::star_of_mass, src/tov.h:205-218, gives)
//   bhx_setup_neutron_star the CLI's neutron-star scene rules (src/main.cpp:472-522)
//
// Host-only scalar work.  The quantities are the reference's; the numerical
// methods are this file's own and more accurate than the reference's, so the
// results agree with the reference's to its discretisation error, not bit for
// bit:
//   * volumes by composite 16-point Gauss-Legendre quadrature on panels split
//     at every kink of the integrand (the reference sums midpoint grids, whose
//     error is ~2e-8 relative for its jet);
//   * the Tolman-Oppenheimer-Volkoff structure in the enthalpy form of
//     Lindblom (1992, ApJ 398, 569): the log-enthalpy h = int dP / (eps + P)
//     is the independent variable, so the surface P = 1e-10 P_c (the
//     reference's surface pressure) is a fixed end point and needs no root
//     search, and the centre is started from the series solution.  The
//     variable of integration is s = sqrt(h_c - h), in which r(s), m(s) are
//     smooth at the centre; fourth-order Runge-Kutta with 8192 steps is
//     converged to ~2e-10 in R (the reference's radius, RK4 in r with the
//     step tied to the pressure scale height, is 4-9e-8 smaller).
#include "../../include/bhx.h"
#include "host_setup.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kG = 6.67430e-11, kC = 2.99792458e8;
constexpr double kSigmaT = 6.6524587321e-29, kMp = 1.67262192369e-27;

// ---- quadrature -----------------------------------------------------------
// 16-point Gauss-Legendre rule on [-1, 1]: nodes by Newton's method on the
// Legendre polynomial P_16 (three-term recurrence), weights 2 / ((1 - x^2) P'^2).
struct GaussRule {
    static constexpr int n = 16;
    double x[n], w[n];
    GaussRule() {
        for (int i = 0; i < n; ++i) {
            double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
            double dp = 1.0;
            for (int it = 0; it < 64; ++it) {
                double p_prev = 1.0, p = z;
                for (int k = 2; k <= n; ++k) {
                    const double p_next = ((2 * k - 1) * z * p - (k - 1) * p_prev) / k;
                    p_prev = p;
                    p = p_next;
                }
                dp = n * (z * p - p_prev) / (z * z - 1.0);
                const double step = p / dp;
                z -= step;
                if (std::fabs(step) < 1e-17) break;
            }
            x[i] = z;
            w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
        }
    }
};

const GaussRule& gauss_rule() {
    static const GaussRule rule;
    return rule;
}

// Integral of f over [lo, hi] on `panels` equal panels.
template <class F>
double integrate(F&& f, double lo, double hi, int panels) {
    const GaussRule& g = gauss_rule();
    const double width = (hi - lo) / panels;
    double total = 0.0;
    for (int p = 0; p < panels; ++p) {
        const double mid = lo + (p + 0.5) * width, half = 0.5 * width;
        double sum = 0.0;
        for (int i = 0; i < GaussRule::n; ++i) sum += g.w[i] * f(mid + half * g.x[i]);
        total += half * sum;
    }
    return total;
}

// The same over [lo, hi] cut at the given interior points (each piece smooth).
template <class F>
double integrate_pieces(F&& f, double lo, double hi, double* cuts, int n_cuts, int panels) {
    std::sort(cuts, cuts + n_cuts);
    double total = 0.0, a = lo;
    for (int i = 0; i <= n_cuts; ++i) {
        const double b = (i < n_cuts) ? std::clamp(cuts[i], lo, hi) : hi;
        if (b > a) total += integrate(f, a, b, panels);
        a = std::max(a, b);
    }
    return total;
}

// ---- accretion flow and jet power ------------------------------------------
// Mdot c^2 [W] of a Novikov-Thorne disc at `eddington` times the Eddington
// rate: L_Edd = 4 pi G M m_p c / sigma_T, efficiency 1 - E at the ISCO
// (what src/disc.h:73-80 and src/constants.h:56-64 give the CLI).
double accretion_power(const bhx_scene& s) {
    const double a = std::clamp(s.spin, -0.9999, 0.9999);
    const int sense = s.disc_sense;
    const double r = bhx::host::isco_radius(a, sense);
    const double u = std::sqrt(r);
    const double E_isco = (r * r - 2.0 * r + sense * a * u) / (r * std::sqrt(r * r - 3.0 * r + 2.0 * sense * a * u));
    const double efficiency = 1.0 - E_isco;
    const double L_edd = 4.0 * kPi * kG * s.M_kg * kMp * kC / kSigmaT;
    return s.eddington * L_edd / efficiency;
}

// Angular velocity of the horizon, a / (2 r_+), in c / r_g.
double horizon_omega(double a) { return a / (2.0 * bhx::host::horizon_outer(a)); }

// Blandford-Znajek power [W] in the form of Tchekhovskoy, Narayan & McKinney
// (2010) with their higher-order spin term: kappa / (4 pi) phi^2 x f(x) Mdot c^2,
// x = Omega_H^2, f = 1 + 1.38 x - 9.2 x^2, kappa = 0.053 (split monopole);
// the reference's jet power (src/jet.h:82).
double jet_power(double a, double mdot_c2, double phi) {
    const double x = horizon_omega(a) * horizon_omega(a);
    const double f = 1.0 + 1.38 * x - 9.2 * x * x;
    return 0.053 / (4.0 * kPi) * phi * phi * x * std::max(0.0, f) * mdot_c2;
}

// Emissivity amplitude at nu_0 = c / 550 nm of a power law nu^-alpha whose band
// [1e9, 1e17] Hz carries the radiated power L_rad [W] over a comoving volume
// V [r_g^3]: 4 pi V j_bol = L_rad, j_bol = j_nu0 nu_0^alpha int nu^-alpha d nu;
// returned per unit r_g of path, as the tracer integrates in r_g.
double emissivity_nu0(double L_rad, double V_rg3, double alpha, double r_g) {
    const double V = V_rg3 * r_g * r_g * r_g;
    if (!(V > 0.0) || !(L_rad > 0.0)) return 0.0;
    const double nu_0 = kC / 550e-9, nu_lo = 1.0e9, nu_hi = 1.0e17;
    const double q = 1.0 - alpha;
    const double band = (std::fabs(q) > 1e-9) ? (std::pow(nu_hi, q) - std::pow(nu_lo, q)) / q
                                              : std::log(nu_hi / nu_lo);
    return L_rad / (4.0 * kPi * V) / (std::pow(nu_0, alpha) * band) * r_g;
}

// ---- the reference's jet geometry (src/jet.h:191-248) ----------------------
// Radius R(z) = R_base (z / z_base)^k, Lorentz factor rising as (z / z_sat)^k
// to gamma_max, transverse profile f(rho / R) of a hollow sheath plus a spine,
// comoving emissivity along z: the beaming and expansion factors
// (gamma / gamma_base)^-(2 + alpha) (R / R_base)^-(3 + alpha) (z / z_base)^zeta
// and the cubic end ramps over 0.5 z_base and 0.15 z_top.
struct SheathJet {
    double z_base, z_top, R_base, k, gamma_max, z_sat, alpha, u_peak, u_wid, spine, spine_wid, zeta;
    double radius(double z) const { return R_base * std::pow(z / z_base, k); }
    double lorentz(double z) const {
        return 1.0 + (gamma_max - 1.0) * std::min(1.0, std::pow(z / z_sat, k));
    }
    double transverse(double u) const {
        const double a = (u - u_peak) / u_wid, b = u / spine_wid;
        return std::exp(-a * a) + spine * std::exp(-b * b);
    }
    static double ramp(double t) {
        t = std::clamp(t, 0.0, 1.0);
        return (3.0 - 2.0 * t) * t * t;
    }
    double axial(double z) const {
        const double beaming = std::pow(lorentz(z) / lorentz(z_base), -(2.0 + alpha));
        const double expansion = std::pow(radius(z) / R_base, -(3.0 + alpha)) * std::pow(z / z_base, zeta);
        return beaming * expansion * ramp((z - z_base) / (0.5 * z_base)) * ramp((z_top - z) / (0.15 * z_top));
    }
    // Emissivity-weighted volume of both jets [r_g^3]: twice the integral
    // over z of axial(z) times the cross-section 2 pi R(z)^2 int_0^3 f(u) u du
    // (the profile's support ends at three sheath widths).
    double volume() const {
        const double section = integrate([&](double u) { return transverse(u) * u; }, 0.0, 3.0, 48);
        double cuts[3] = {1.5 * z_base, z_sat, 0.85 * z_top};   // ramp ends, saturation
        const double along = integrate_pieces([&](double z) { return axial(z) * radius(z) * radius(z); },
                                              z_base, z_top, cuts, 3, 24);
        return 2.0 * 2.0 * kPi * section * along;
    }
};

// ---- the conical jet --------------------------------------------------------
// Emissivity-weighted volume of both cones [r_g^3]: profile (r0/r)^2 x taper
// in mu = |cos theta| x end ramps, the taper and the ramps quintic steps S with
// int_0^1 S = 1/2.  Angular part: int taper d mu = (1 - cos 0.8 th) + (cos 0.8 th
// - cos th) / 2; radial part: r0^2 times the integral of the two ramps.
double cone_volume(double th, double r0, double r1) {
    const double c_edge = std::cos(th), c_core = std::cos(0.8 * th);
    const double angular = (1.0 - c_core) + 0.5 * (c_core - c_edge);
    auto step = [](double t) {
        t = std::clamp(t, 0.0, 1.0);
        return t * t * t * (10.0 + t * (6.0 * t - 15.0));
    };
    double cuts[2] = {1.5 * r0, 0.85 * r1};
    const double radial = r0 * r0 * integrate_pieces([&](double r) {
        return step((r - r0) / (0.5 * r0)) * step((r1 - r) / (0.15 * r1)); }, r0, r1, cuts, 2, 8);
    return 2.0 * 2.0 * kPi * angular * radial;
}

// ---- TOV in the enthalpy form ----------------------------------------------
// Polytrope p = K eps^Gamma in geometrised units (G = c = 1, lengths in metres):
// eps = G rho / c^2 the mass-energy density, p = G P / c^4.  Its log-enthalpy
// h(eps) = int dp / (eps + p) = Gamma / (Gamma - 1) ln(1 + K eps^(Gamma - 1)),
// inverted in closed form.
struct EnthalpyPolytrope {
    double K, Gam;
    EnthalpyPolytrope(double K_si, double Gamma)
        : K(K_si * std::pow(kG, 1.0 - Gamma) * std::pow(kC, 2.0 * Gamma - 4.0)), Gam(Gamma) {}
    double pressure(double eps) const { return K * std::pow(eps, Gam); }
    double enthalpy(double eps) const { return Gam / (Gam - 1.0) * std::log1p(K * std::pow(eps, Gam - 1.0)); }
    double energy_density(double h) const {
        const double x = std::expm1(h * (Gam - 1.0) / Gam);
        return (x > 0.0) ? std::pow(x / K, 1.0 / (Gam - 1.0)) : 0.0;
    }
};

struct Structure { double R_m = 0.0, M_kg = 0.0; bool ok = false; };

// One star from its central density rho_c [kg m^-3].  With h = h_c - s^2:
//   dr/ds = 2 s r (r - 2 m) / (m + 4 pi r^3 p),   dm/ds = 4 pi r^2 eps dr/ds,
// from the centre (series of Lindblom 1992, eqs. 7-8, at s0 = 1e-4 s_surface)
// to the surface p = 1e-10 p_c.
Structure tov_structure(const EnthalpyPolytrope& eos, double rho_c) {
    Structure out;
    const double eps_c = kG * rho_c / (kC * kC);
    const double p_c = eos.pressure(eps_c);
    if (!(p_c > 0.0) || !std::isfinite(p_c)) return out;
    const double h_c = eos.enthalpy(eps_c);
    const double eps_surface = std::pow(1e-10 * p_c / eos.K, 1.0 / eos.Gam);
    const double s_surface = std::sqrt(h_c - eos.enthalpy(eps_surface));

    double s = 1e-4 * s_surface;
    const double dh = s * s;
    const double deps_dh = (eps_c + p_c) * eps_c / (eos.Gam * p_c);   // d eps / dh at the centre
    double r = std::sqrt(3.0 * dh / (2.0 * kPi * (eps_c + 3.0 * p_c))) *
               (1.0 - 0.25 * (eps_c - 3.0 * p_c - 0.6 * deps_dh) * dh / (eps_c + 3.0 * p_c));
    double m = 4.0 * kPi / 3.0 * eps_c * r * r * r * (1.0 - 0.6 * deps_dh * dh / eps_c);

    bool bad = false;
    auto slope = [&](double s_at, double r_at, double m_at, double& dr, double& dm) {
        const double eps = eos.energy_density(h_c - s_at * s_at);
        const double p = eos.pressure(eps);
        const double weight = m_at + 4.0 * kPi * r_at * r_at * r_at * p;
        const double rim = r_at - 2.0 * m_at;
        if (!(weight > 0.0) || !(rim > 0.0)) { bad = true; dr = dm = 0.0; return; }
        dr = 2.0 * s_at * r_at * rim / weight;
        dm = 4.0 * kPi * r_at * r_at * eps * dr;
    };
    constexpr int kSteps = 8192;
    const double H = (s_surface - s) / kSteps;
    for (int i = 0; i < kSteps && !bad; ++i) {
        double r1, m1, r2, m2, r3, m3, r4, m4;
        slope(s, r, m, r1, m1);
        slope(s + 0.5 * H, r + 0.5 * H * r1, m + 0.5 * H * m1, r2, m2);
        slope(s + 0.5 * H, r + 0.5 * H * r2, m + 0.5 * H * m2, r3, m3);
        slope(s + H, r + H * r3, m + H * m3, r4, m4);
        r += H / 6.0 * (r1 + 2.0 * (r2 + r3) + r4);
        m += H / 6.0 * (m1 + 2.0 * (m2 + m3) + m4);
        s = s_surface - (kSteps - 1 - i) * H;
    }
    if (bad || !std::isfinite(r) || !std::isfinite(m) || !(r > 2.0 * m)) return out;
    out.R_m = r;
    out.M_kg = m * kC * kC / kG;
    out.ok = out.M_kg > 0.0;
    return out;
}

// The star of gravitational mass M_target on the stable branch of central
// densities in [rho_lo, rho_hi] (the reference's search range): the mass
// maximum by golden-section search in ln rho_c, then the root of
// M(rho_c) = M_target below it by the Illinois variant of regula falsi.
int stable_star(const EnthalpyPolytrope& eos, double M_target, Structure& star,
                double rho_lo = 2e17, double rho_hi = 5e18) {
    auto mass = [&](double x) {
        const Structure st = tov_structure(eos, std::exp(x));
        return st.ok ? st.M_kg : 0.0;
    };
    const double invphi = 0.5 * (std::sqrt(5.0) - 1.0);
    double a = std::log(rho_lo), b = std::log(rho_hi);
    double x1 = b - invphi * (b - a), x2 = a + invphi * (b - a);
    double f1 = mass(x1), f2 = mass(x2);
    while (b - a > 1e-6) {
        if (f1 < f2) { a = x1; x1 = x2; f1 = f2; x2 = a + invphi * (b - a); f2 = mass(x2); }
        else { b = x2; x2 = x1; f2 = f1; x1 = b - invphi * (b - a); f1 = mass(x1); }
    }
    double x_hi = (f1 > f2) ? x1 : x2;
    double g_hi = std::max(f1, f2) - M_target;
    const double f_end = mass(std::log(rho_hi));
    if (f_end - M_target > g_hi) { x_hi = std::log(rho_hi); g_hi = f_end - M_target; }
    if (g_hi < 0.0) return BHX_E_INVALID;                 // heavier than the maximum mass
    double x_lo = std::log(rho_lo), g_lo = mass(x_lo) - M_target;
    if (g_lo > 0.0) return BHX_E_INVALID;                 // lighter than the range's lightest star
    int side = 0;
    for (int it = 0; it < 200 && x_hi - x_lo > 1e-14 * std::fabs(x_hi); ++it) {
        const double x = (x_lo * g_hi - x_hi * g_lo) / (g_hi - g_lo);
        const double g = mass(x) - M_target;
        if (g == 0.0) { x_lo = x_hi = x; break; }
        if (g < 0.0) {
            x_lo = x; g_lo = g;
            if (side == -1) g_hi *= 0.5;
            side = -1;
        } else {
            x_hi = x; g_hi = g;
            if (side == +1) g_lo *= 0.5;
            side = +1;
        }
    }
    star = tov_structure(eos, std::exp(0.5 * (x_lo + x_hi)));
    return star.ok ? BHX_OK : BHX_E_INVALID;
}

}  // namespace

extern "C" {

int bhx_setup_jet(bhx_scene* s, double phi_flux, double eps_rad) {
    if (!s || !(phi_flux >= 0.0) || !(eps_rad >= 0.0)) return BHX_E_INVALID;
    if (!s->jet_enabled) return BHX_OK;
    if (s->ns_enabled || !s->disc_enabled) {  // no horizon / no accretion flow: no jet
        s->jet_enabled = 0;
        return BHX_OK;
    }
    s->jet_z_sat = std::min(s->jet_z_sat, 0.7 * s->jet_z_top);   // the CLI's cap (src/main.cpp:415)
    const SheathJet jet{s->jet_z_base, s->jet_z_top, s->jet_R_base, s->jet_k, s->jet_gamma_max, s->jet_z_sat,
                        s->jet_alpha, s->jet_u_peak, s->jet_u_wid, s->jet_spine, s->jet_spine_wid, s->jet_zeta};
    const double P_jet = jet_power(s->spin, accretion_power(*s), phi_flux);
    const double R_top = jet.radius(jet.z_top);
    s->jet_a = s->spin;
    s->jet_Omega_F = 0.5 * horizon_omega(s->spin);       // field lines at half the horizon's rate
    s->jet_r_bound = 1.05 * std::sqrt(jet.z_top * jet.z_top + R_top * R_top);
    s->jet_j_nu0 = emissivity_nu0(eps_rad * P_jet, jet.volume(), jet.alpha, kG * s->M_kg / (kC * kC));
    if (!(P_jet > 0.0)) s->jet_enabled = 0;  // a = 0: nothing to tap
    return BHX_OK;
}

int bhx_setup_conical_jet(bhx_scene* s, double phi_flux, double eps_rad) {
    if (!s || !(phi_flux >= 0.0) || !(eps_rad >= 0.0)) return BHX_E_INVALID;
    if (!s->jet_enabled) return BHX_OK;
    if (!(s->jet_cone_angle > 0.0) || !(s->jet_cone_r0 > 0.0) || !(s->jet_cone_r1 > s->jet_cone_r0))
        return BHX_E_INVALID;
    if (s->ns_enabled || !s->disc_enabled) {
        s->jet_enabled = 0;
        return BHX_OK;
    }
    s->jet_model = 1;
    const double P_jet = jet_power(s->spin, accretion_power(*s), phi_flux);
    const double V = cone_volume(s->jet_cone_angle, s->jet_cone_r0, s->jet_cone_r1);
    s->jet_a = s->spin;
    s->jet_Omega_F = 0.5 * horizon_omega(s->spin);
    s->jet_j_nu0 = emissivity_nu0(eps_rad * P_jet, V, s->jet_alpha, kG * s->M_kg / (kC * kC));
    s->jet_r_bound = 1.05 * s->jet_cone_r1;
    if (!(P_jet > 0.0)) s->jet_enabled = 0;
    return BHX_OK;
}

int bhx_polytrope_star(double K_si, double Gamma, double M_kg, double* R_m, double* M_out) {
    if (!(K_si > 0.0) || !(Gamma > 1.0) || !(M_kg > 0.0) || !R_m || !M_out) return BHX_E_INVALID;
    Structure st;
    const int rc = stable_star(EnthalpyPolytrope(K_si, Gamma), M_kg, st);
    *R_m = st.R_m;
    *M_out = st.M_kg;
    return rc;
}

int bhx_uniform_sky(int32_t count, uint64_t seed, double* x, double* y, double* z, double* temperature,
                    double* irradiance) {
    if (count < 0 || (count > 0 && (!x || !y || !z || !temperature || !irradiance))) return BHX_E_INVALID;
    bhx::host::Catalogue cat;
    bhx::host::uniform_sky(count, seed, cat);
    std::copy(cat.x.begin(), cat.x.end(), x);
    std::copy(cat.y.begin(), cat.y.end(), y);
    std::copy(cat.z.begin(), cat.z.end(), z);
    std::copy(cat.T.begin(), cat.T.end(), temperature);
    std::copy(cat.irr.begin(), cat.irr.end(), irradiance);
    return BHX_OK;
}

int bhx_setup_neutron_star(bhx_scene* s, double R_m, double M_kg) {
    if (!s || !(R_m > 0.0) || !(M_kg > 0.0)) return BHX_E_INVALID;
    const double r_g = kG * M_kg / (kC * kC);
    s->M_kg = M_kg;
    s->spin = 0.0;              // Birkhoff: the exterior is Schwarzschild
    s->disc_enabled = 0;
    s->jet_enabled = 0;
    s->ns_enabled = 1;
    s->ns_M_kg = M_kg;
    s->ns_R_m = R_m;
    s->ns_R = R_m / r_g;        // NeutronStar::set_from (neutronstar.h:69-73)
    s->ns_Omega = 2.0 * kPi * s->ns_spin_hz * r_g / kC;   // set_spin (neutronstar.h:75-79)
    return BHX_OK;
}

}  // extern "C"
