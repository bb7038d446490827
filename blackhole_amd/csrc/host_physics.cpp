// host_physics.cpp - host solvers that prepare a scene before tracing.
//
//   bhx_setup_jet        the CLI's setup_jet (src/main.cpp:399-423) with
//                        Jet::configure (src/jet.h:261-311): Blandford-Znajek
//                        power -> radiated fraction -> emissivity normalisation
//   bhx_polytrope_star   ns::star_of_mass (src/tov.h:205-218) for P = K rho^Gamma
//                        (src/eos.h:153-167): RK4 TOV integration with surface
//                        bisection (src/tov.h:65-177)
//   bhx_setup_neutron_star  the CLI's NS scene rules (src/main.cpp:472-522)
//
// Host-only scalar work (milliseconds); plain IEEE double in the reference's
// operation order, so the results equal the reference's to the last bit on the
// same libm.
#include "../../include/bhx.h"
#include "host_setup.h"

#include <algorithm>
#include <cmath>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kG = 6.67430e-11, kC = 2.99792458e8;
constexpr double kSigmaT = 6.6524587321e-29, kMp = 1.67262192369e-27, kMsun = 1.98847e30;

// ---- jets ----------------------------------------------------------------
struct JetCfg {
    double z_base, z_top, R_base, k, gamma_max, z_sat, alpha, u_peak, u_wid, spine, spine_wid, zeta;
    double radius(double z) const { return R_base * std::pow(std::max(z, 1e-6) / z_base, k); }
    double gamma_at(double z) const {
        const double t = std::min(1.0, std::pow(std::max(z, 1e-6) / z_sat, k));
        return 1.0 + (gamma_max - 1.0) * t;
    }
    double profile(double u) const {
        const double s = (u - u_peak) / u_wid;
        const double t = u / spine_wid;
        return std::exp(-s * s) + spine * std::exp(-t * t);
    }
};

double smooth01(double t) {
    t = std::clamp(t, 0.0, 1.0);
    return t * t * (3.0 - 2.0 * t);
}

double smooth_c2(double t) {   // the conical jet's quintic step
    t = std::clamp(t, 0.0, 1.0);
    return t * t * t * (10.0 + t * (6.0 * t - 15.0));
}

double horizon_omega(double a) {  // kerr.h:280-283
    const double rp = 1.0 + std::sqrt(std::max(0.0, 1.0 - a * a));
    return (rp > 0.0) ? a / (2.0 * rp) : 0.0;
}

double circ_energy(double a, double r, int sg) {  // kerr.h:232-237
    const double sr = std::sqrt(r);
    const double num = r * r - 2.0 * r + sg * a * sr;
    const double den = r * std::sqrt(r * r - 3.0 * r + 2.0 * sg * a * sr);
    return num / den;
}

// ---- TOV -------------------------------------------------------------------
struct Tov {
    double K, Gam;
    double pressure(double rho) const { return (rho > 0.0) ? K * std::pow(rho, Gam) : 0.0; }
    double density(double P) const { return (P > 0.0) ? std::pow(P / K, 1.0 / Gam) : 0.0; }
};

struct Y4 { double m, P, phi, mb; };

struct Star { double R = 0.0, M = 0.0, rho_c = 0.0; bool ok = false; };

// One star from its central density (tov.h:65-177): RK4 in r, step from the
// pressure scale height, the surface P = 1e-10 P_c located by bisection on the
// step length.
Star solve_tov(const Tov& eos, double rho_c) {
    const double c2 = kC * kC;
    Star s;
    s.rho_c = rho_c;
    const double P_c = eos.pressure(rho_c);
    if (!(P_c > 0.0)) return s;
    const double P_stop = P_c * 1e-10;
    auto deriv = [&](double r, const Y4& y) -> Y4 {
        Y4 d{};
        if (r <= 0.0) return d;
        const double rho = eos.density(y.P);
        const double f = 1.0 - 2.0 * kG * y.m / (r * c2);
        d.m = 4.0 * kPi * r * r * rho;
        if (f <= 1e-12) {
            d.P = -1e300;
            d.phi = 0.0;
            d.mb = 0.0;
            return d;
        }
        d.P = -kG * (rho + y.P / c2) * (y.m + 4.0 * kPi * r * r * r * y.P / c2) / (r * r * f);
        d.phi = -d.P / (rho * c2 + y.P);
        d.mb = 4.0 * kPi * r * r * rho / std::sqrt(f);
        return d;
    };
    auto add = [](const Y4& a, const Y4& b, double h) {
        return Y4{a.m + h * b.m, a.P + h * b.P, a.phi + h * b.phi, a.mb + h * b.mb};
    };
    double r = 1e-3;
    Y4 y{};
    y.m = (4.0 / 3.0) * kPi * rho_c * r * r * r;
    y.P = P_c - (2.0 * kPi * kG / 3.0) * (rho_c + P_c / c2) * (rho_c + 3.0 * P_c / c2) * r * r;
    for (int step = 0; step < 2000000; ++step) {
        const Y4 d0 = deriv(r, y);
        if (d0.P < -1e299) return s;
        const double scale = (d0.P != 0.0) ? std::fabs(y.P / d0.P) : 1e30;
        const double dr = std::clamp(0.02 * scale, 0.1, 200.0);
        const Y4 k1 = d0;
        const Y4 k2 = deriv(r + 0.5 * dr, add(y, k1, 0.5 * dr));
        const Y4 k3 = deriv(r + 0.5 * dr, add(y, k2, 0.5 * dr));
        const Y4 k4 = deriv(r + dr, add(y, k3, dr));
        Y4 yn{};
        yn.m = y.m + dr / 6.0 * (k1.m + 2 * k2.m + 2 * k3.m + k4.m);
        yn.P = y.P + dr / 6.0 * (k1.P + 2 * k2.P + 2 * k3.P + k4.P);
        yn.phi = y.phi + dr / 6.0 * (k1.phi + 2 * k2.phi + 2 * k3.phi + k4.phi);
        yn.mb = y.mb + dr / 6.0 * (k1.mb + 2 * k2.mb + 2 * k3.mb + k4.mb);
        if (yn.P <= P_stop || !std::isfinite(yn.P)) {
            double lo = 0.0, hi = dr;
            for (int k = 0; k < 80; ++k) {
                const double mid = 0.5 * (lo + hi);
                const Y4 a1 = deriv(r, y);
                const Y4 a2 = deriv(r + 0.5 * mid, add(y, a1, 0.5 * mid));
                const Y4 a3 = deriv(r + 0.5 * mid, add(y, a2, 0.5 * mid));
                const Y4 a4 = deriv(r + mid, add(y, a3, mid));
                const double Pm = y.P + mid / 6.0 * (a1.P + 2 * a2.P + 2 * a3.P + a4.P);
                if (Pm <= P_stop || !std::isfinite(Pm)) hi = mid; else lo = mid;
            }
            const double h = 0.5 * (lo + hi);
            const Y4 a1 = deriv(r, y);
            const Y4 a2 = deriv(r + 0.5 * h, add(y, a1, 0.5 * h));
            const Y4 a3 = deriv(r + 0.5 * h, add(y, a2, 0.5 * h));
            const Y4 a4 = deriv(r + h, add(y, a3, h));
            s.R = r + h;
            s.M = y.m + h / 6.0 * (a1.m + 2 * a2.m + 2 * a3.m + a4.m);
            s.ok = s.R > 0.0 && s.M > 0.0;
            break;
        }
        y = yn;
        r += dr;
    }
    return s;
}

// Stable-branch star of a given mass (tov.h:188-218).
Star star_of_mass(const Tov& eos, double M_target, double rho_lo = 2e17, double rho_hi = 5e18) {
    Star best;
    for (int i = 0; i < 40; ++i) {
        const double f = static_cast<double>(i) / (40 - 1);
        const Star s = solve_tov(eos, rho_lo * std::pow(rho_hi / rho_lo, f));
        if (!s.ok) continue;
        if (s.M > best.M) best = s;
    }
    if (best.M < M_target) return Star{};
    double lo = rho_lo, hi = best.rho_c;
    for (int i = 0; i < 80; ++i) {
        const double mid = std::sqrt(lo * hi);
        const Star s = solve_tov(eos, mid);
        if (!s.ok || s.M < M_target) lo = mid; else hi = mid;
    }
    return solve_tov(eos, std::sqrt(lo * hi));
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
    JetCfg j{s->jet_z_base, s->jet_z_top, s->jet_R_base, s->jet_k, s->jet_gamma_max, s->jet_z_sat,
             s->jet_alpha, s->jet_u_peak, s->jet_u_wid, s->jet_spine, s->jet_spine_wid, s->jet_zeta};
    // NovikovThorneDisc ctor: efficiency and Mdot (disc.h:73-80, constants.h:56-64)
    const double a_d = std::clamp(s->spin, -0.9999, 0.9999);
    const double r_isco = bhx::host::isco_radius(a_d, s->disc_sense);
    const double eta = 1.0 - circ_energy(a_d, r_isco, s->disc_sense);
    const double L_edd = 4.0 * kPi * kG * s->M_kg * kMp * kC / kSigmaT;
    const double mdot = s->eddington * (L_edd / (eta * kC * kC));
    const double mdot_c2 = mdot * kC * kC;
    const double r_g_m = kG * s->M_kg / (kC * kC);
    j.z_sat = std::min(j.z_sat, 0.7 * j.z_top);                    // main.cpp:415
    // Jet::configure (jet.h:261-311)
    const double spin = s->spin;
    const double Omega_F = 0.5 * horizon_omega(spin);
    const double OmH = horizon_omega(spin);
    const double x = OmH * OmH;
    const double f = 1.0 + 1.38 * x - 9.2 * x * x;
    const double phi = phi_flux, kappa = 0.053;                      // jet.h:82
    const double P_jet = (kappa / (4.0 * kPi)) * phi * phi * x * std::max(0.0, f) * mdot_c2;
    const double L_rad = eps_rad * P_jet;
    const double r_bound = std::sqrt(j.z_top * j.z_top + j.radius(j.z_top) * j.radius(j.z_top)) * 1.05;
    double C_f = 0.0;
    {
        const int N = 2000;
        const double du = 3.0 / N;
        for (int i = 0; i < N; ++i) {
            const double u = (i + 0.5) * du;
            C_f += j.profile(u) * u * du;
        }
    }
    double V = 0.0;
    {
        const int N = 4000;
        const double dz = (j.z_top - j.z_base) / N;
        for (int i = 0; i < N; ++i) {
            const double z = j.z_base + (i + 0.5) * dz;
            const double R = j.radius(z);
            const double gz = std::pow(j.gamma_at(z) / j.gamma_at(j.z_base), -(2.0 + j.alpha)) *
                              std::pow(R / j.R_base, -(3.0 + j.alpha)) * std::pow(z / j.z_base, j.zeta);
            const double ramp = smooth01((z - j.z_base) / (0.5 * j.z_base)) *
                                smooth01((j.z_top - z) / (0.15 * j.z_top));
            V += gz * ramp * R * R * dz;
        }
        V *= 2.0 * (2.0 * kPi) * C_f;
    }
    const double V_m3 = V * r_g_m * r_g_m * r_g_m;
    double j_nu0 = 0.0;
    if (V_m3 > 0.0 && L_rad > 0.0) {
        const double A = L_rad / (4.0 * kPi * V_m3);
        const double nu_0 = kC / 550e-9, nu_lo = 1.0e9, nu_hi = 1.0e17;
        const double p = 1.0 - j.alpha;
        const double band = (std::fabs(p) > 1e-9) ? (std::pow(nu_hi, p) - std::pow(nu_lo, p)) / p
                                                  : std::log(nu_hi / nu_lo);
        j_nu0 = A / (std::pow(nu_0, j.alpha) * band) * r_g_m;
    }
    s->jet_z_sat = j.z_sat;
    s->jet_a = spin;
    s->jet_Omega_F = Omega_F;
    s->jet_j_nu0 = j_nu0;
    s->jet_r_bound = r_bound;
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
    // accretion rate and Blandford-Znajek power as for the reference's jets
    const double a_d = std::clamp(s->spin, -0.9999, 0.9999);
    const double r_isco = bhx::host::isco_radius(a_d, s->disc_sense);
    const double eta = 1.0 - circ_energy(a_d, r_isco, s->disc_sense);
    const double L_edd = 4.0 * kPi * kG * s->M_kg * kMp * kC / kSigmaT;
    const double mdot_c2 = s->eddington * (L_edd / (eta * kC * kC)) * kC * kC;
    const double r_g_m = kG * s->M_kg / (kC * kC);
    const double OmH = horizon_omega(s->spin);
    const double x = OmH * OmH;
    const double f = 1.0 + 1.38 * x - 9.2 * x * x;
    const double P_jet = (0.053 / (4.0 * kPi)) * phi_flux * phi_flux * x * std::max(0.0, f) * mdot_c2;
    const double L_rad = eps_rad * P_jet;
    // emitting volume of both cones, weighted by the profile (r0/r)^2 x taper x ramps
    const double th_j = s->jet_cone_angle, r0 = s->jet_cone_r0, r1 = s->jet_cone_r1;
    // angular part: integral of the taper over sin psi d psi = d(cos psi),
    // the taper being smooth_c2((mu - cos th_j) / (cos 0.8 th_j - cos th_j))
    double ang = 0.0;
    {
        const int N = 400;
        const double c0 = std::cos(th_j), c1 = std::cos(0.8 * th_j);
        ang = 1.0 - c1;   // mu in [cos 0.8 th_j, 1]: taper 1
        const double dmu = (c1 - c0) / N;
        for (int i = 0; i < N; ++i) ang += smooth_c2((i + 0.5) / N) * dmu;
    }
    double rad = 0.0;
    {
        const int N = 8000;
        const double dr = (r1 - r0) / N;
        for (int i = 0; i < N; ++i) {
            const double r = r0 + (i + 0.5) * dr;
            const double q = r0 / r;
            rad += q * q * smooth_c2((r - r0) / (0.5 * r0)) * smooth_c2((r1 - r) / (0.15 * r1)) * r * r * dr;
        }
    }
    const double V = 2.0 * (2.0 * kPi) * ang * rad;
    const double V_m3 = V * r_g_m * r_g_m * r_g_m;
    double j_nu0 = 0.0;
    if (V_m3 > 0.0 && L_rad > 0.0) {
        const double A = L_rad / (4.0 * kPi * V_m3);
        const double nu_0 = kC / 550e-9, nu_lo = 1.0e9, nu_hi = 1.0e17;
        const double p = 1.0 - s->jet_alpha;
        const double band = (std::fabs(p) > 1e-9) ? (std::pow(nu_hi, p) - std::pow(nu_lo, p)) / p
                                                  : std::log(nu_hi / nu_lo);
        j_nu0 = A / (std::pow(nu_0, s->jet_alpha) * band) * r_g_m;
    }
    s->jet_a = s->spin;
    s->jet_Omega_F = 0.5 * OmH;
    s->jet_j_nu0 = j_nu0;
    s->jet_r_bound = 1.05 * r1;
    if (!(P_jet > 0.0)) s->jet_enabled = 0;
    return BHX_OK;
}

int bhx_polytrope_star(double K_si, double Gamma, double M_kg, double* R_m, double* M_out) {
    if (!(K_si > 0.0) || !(Gamma > 1.0) || !(M_kg > 0.0) || !R_m || !M_out) return BHX_E_INVALID;
    const Star st = star_of_mass(Tov{K_si, Gamma}, M_kg);
    *R_m = st.R;
    *M_out = st.M;
    return st.ok ? BHX_OK : BHX_E_INVALID;
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
