"""Host scene preparation in the C-ABI library (jets, TOV, neutron-star scene)
against the values the reference build wrote into the golden fixtures, plus
the closed forms and independent integrations that pin the TOV solver where
no fixture reaches.  CPU only: these entry points make no GPU call.

Tolerances against the reference's values are its own discretisation errors:
  j_nu0   3e-8 relative: the reference sums its jet volume on midpoint grids
          (2000 and 4000 cells, src/jet.h:271-300), 0.7-1.7e-8 from the
          integrals, which the library evaluates by Gauss-Legendre quadrature
          (checked against scipy's adaptive quadrature to 1e-11 below);
  R       1e-7 relative: the reference integrates the TOV equations by RK4 in
          r with the step tied to the pressure scale height (src/tov.h:65-177),
          4-9e-8 from the converged radius; the library's enthalpy-form
          integration agrees with scipy's DOP853 in r at rtol 1e-12 to 1e-9;
  M       1e-12 relative: both root-find the central density for the mass;
  r_bound, Omega_F, z_sat: closed forms, equal.
"""
import math

import pytest
from scipy import integrate, optimize

from blackhole_amd import scene as S
from blackhole_amd.native import BhxError
from test_oracle_golden import META

G, C_LIGHT = 6.67430e-11, 2.99792458e8
SIGMA_T, M_P = 6.6524587321e-29, 1.67262192369e-27


@pytest.mark.parametrize("case", META["jets"], ids=lambda c: f"a{c['spin']}_edd{c['eddington']}")
def test_setup_jet_matches_reference(case):
    sc = S.default_scene()
    sc.M_kg, sc.spin, sc.eddington = 1e9 * S.M_SUN, case["spin"], case["eddington"]
    sc.disc_r_out, sc.jet_z_top, sc.jet_enabled = 20.0, case["z_top"], 1
    S.setup_jet(sc)
    assert sc.jet_enabled == 1
    assert sc.jet_a == case["spin"]
    for k in ("r_bound", "Omega_F", "z_sat"):
        assert getattr(sc, "jet_" + k) == case[k], k
    assert sc.jet_j_nu0 == pytest.approx(case["j_nu0"], rel=3e-8)


def test_setup_jet_golden_scene_fields():
    """The quasar scene's stored jet fields came from the reference's setup."""
    d = META["scenes"]["quasar_jets"]["scene"]
    sc = S.default_scene()
    for k in ("M_kg", "spin", "eddington", "disc_r_out", "disc_sense", "jet_z_top", "jet_gamma_max"):
        setattr(sc, k, d[k])
    sc.jet_enabled = 1
    S.setup_jet(sc)
    for k in ("jet_z_sat", "jet_a", "jet_Omega_F", "jet_r_bound"):
        assert getattr(sc, k) == d[k], k
    assert sc.jet_j_nu0 == pytest.approx(d["jet_j_nu0"], rel=3e-8)


def _sheath_volume(sc):
    """The reference jet's emissivity-weighted volume (src/jet.h:261-300) by
    scipy's adaptive quadrature, in units of r_g^3."""
    zb, zt, Rb, k = sc.jet_z_base, sc.jet_z_top, sc.jet_R_base, sc.jet_k
    gm, zs, al, ze = sc.jet_gamma_max, sc.jet_z_sat, sc.jet_alpha, sc.jet_zeta

    def prof(u):
        return math.exp(-((u - sc.jet_u_peak) / sc.jet_u_wid) ** 2) + sc.jet_spine * math.exp(-(u / sc.jet_spine_wid) ** 2)

    def ramp(t):
        t = min(max(t, 0.0), 1.0)
        return t * t * (3 - 2 * t)

    R = lambda z: Rb * (z / zb) ** k
    gam = lambda z: 1 + (gm - 1) * min(1.0, (z / zs) ** k)

    def axial(z):
        return ((gam(z) / gam(zb)) ** (-(2 + al)) * (R(z) / Rb) ** (-(3 + al)) * (z / zb) ** ze
                * ramp((z - zb) / (0.5 * zb)) * ramp((zt - z) / (0.15 * zt)) * R(z) ** 2)

    cf, _ = integrate.quad(lambda u: prof(u) * u, 0, 3, epsabs=0, epsrel=1e-13, limit=400)
    L, _ = integrate.quad(axial, zb, zt, points=[1.5 * zb, zs, 0.85 * zt], epsabs=0, epsrel=1e-13, limit=400)
    return 4 * math.pi * cf * L


def test_setup_jet_volume_against_quadrature():
    """4 pi j_nu0 nu0^alpha band V r_g^2 = eps_rad P_BZ with V from scipy: the
    library's quadrature is exact where the reference's midpoint sums are not."""
    import oracle
    for case in META["jets"]:
        sc = S.default_scene()
        sc.M_kg, sc.spin, sc.eddington = 1e9 * S.M_SUN, case["spin"], case["eddington"]
        sc.disc_r_out, sc.jet_z_top, sc.jet_enabled = 20.0, case["z_top"], 1
        S.setup_jet(sc)
        r_g = G * sc.M_kg / C_LIGHT ** 2
        radiated = 4 * math.pi * sc.jet_j_nu0 * oracle.jet_band_factor(sc.jet_alpha) * _sheath_volume(sc) * r_g ** 2
        # the same power from the reference's own normalisation (golden j_nu0)
        ref = 4 * math.pi * case["j_nu0"] * oracle.jet_band_factor(sc.jet_alpha) * _sheath_volume(sc) * r_g ** 2
        assert radiated == pytest.approx(ref, rel=3e-8)
        # eps_rad x P_BZ with P_BZ = kappa/(4 pi) phi^2 Omega_H^2 f Mdot c^2
        a = sc.spin
        r_isco = oracle.lib().oracle_isco(a, 1)
        sr = math.sqrt(r_isco)
        E = (r_isco ** 2 - 2 * r_isco + a * sr) / (r_isco * math.sqrt(r_isco ** 2 - 3 * r_isco + 2 * a * sr))
        L_edd = 4 * math.pi * G * sc.M_kg * M_P * C_LIGHT / SIGMA_T
        om = a / (2.0 * (1.0 + math.sqrt(1.0 - a * a)))
        x = om * om
        P_bz = 0.053 / (4 * math.pi) * 50.0 ** 2 * x * (1 + 1.38 * x - 9.2 * x * x) * sc.eddington * L_edd / (1 - E)
        assert radiated == pytest.approx(0.02 * P_bz, rel=1e-11)


def test_setup_jet_switch_offs():
    sc = S.default_scene()
    sc.jet_enabled, sc.spin = 1, 0.0          # no rotational energy to tap
    S.setup_jet(sc)
    assert sc.jet_enabled == 0
    sc = S.default_scene()
    sc.jet_enabled, sc.disc_enabled = 1, 0    # nothing threads the horizon
    S.setup_jet(sc)
    assert sc.jet_enabled == 0
    sc = S.default_scene()
    sc.jet_enabled, sc.ns_enabled = 1, 1      # no horizon
    S.setup_jet(sc)
    assert sc.jet_enabled == 0


def test_jet_power_scales_with_accretion_rate():
    """P_jet and so j_nu0 are linear in Mdot; Omega_F = Omega_H/2 = a/(4 r_+)."""
    out = []
    for edd in (0.01, 0.1):
        sc = S.default_scene()
        sc.M_kg, sc.spin, sc.eddington, sc.jet_enabled = 1e9 * S.M_SUN, 0.9, edd, 1
        S.setup_jet(sc)
        out.append(sc.jet_j_nu0)
        assert sc.jet_Omega_F == pytest.approx(0.9 / (4.0 * (1.0 + math.sqrt(1.0 - 0.81))), rel=1e-15)
    assert out[1] / out[0] == pytest.approx(10.0, rel=1e-12)


def test_polytrope_stars_match_reference():
    t = META["tov"]
    for st in t["stars"]:
        R, M = S.polytrope_star(t["K_si"], t["Gamma"], st["M_target"] * S.M_SUN)
        assert st["ok"]
        assert R == pytest.approx(st["R_m"], rel=1e-7)
        assert M == pytest.approx(st["M_kg"], rel=1e-12)


def _tov_in_r(K_si, gamma, rho_c):
    """TOV structure integrated in r (the reference's own formulation,
    src/tov.h:1-30) by scipy's DOP853 at rtol 1e-12, to the surface
    P = 1e-10 P_c: (R [m], M [kg])."""
    c2 = C_LIGHT ** 2
    Pc = K_si * rho_c ** gamma

    def rhs(r, y):
        m, P = y
        rho = (max(P, 0.0) / K_si) ** (1 / gamma)
        f = 1 - 2 * G * m / (r * c2)
        return [4 * math.pi * r * r * rho,
                -G * (rho + P / c2) * (m + 4 * math.pi * r ** 3 * P / c2) / (r * r * f)]

    def surface(r, y):
        return y[1] - 1e-10 * Pc
    surface.terminal = True
    r0 = 1e-2
    y0 = [4 / 3 * math.pi * rho_c * r0 ** 3,
          Pc - 2 * math.pi * G / 3 * (rho_c + Pc / c2) * (rho_c + 3 * Pc / c2) * r0 ** 2]
    sol = integrate.solve_ivp(rhs, (r0, 1e6), y0, method="DOP853", rtol=1e-12, atol=[1e-30, 1e-30 * Pc],
                              events=surface)
    return sol.t_events[0][0], sol.y_events[0][0][0]


def test_polytrope_against_integration_in_r():
    """The enthalpy-form structure equals the TOV equations integrated in r:
    for each golden mass, the central density that scipy's integration needs
    for that mass gives the library's radius to 1e-9."""
    t = META["tov"]
    for st in t["stars"][::2]:
        M_t = st["M_target"] * S.M_SUN
        R, M = S.polytrope_star(t["K_si"], t["Gamma"], M_t)
        lr = optimize.brentq(lambda x: _tov_in_r(t["K_si"], t["Gamma"], math.exp(x))[1] - M_t,
                             math.log(2e17), math.log(1.5e18), xtol=1e-13)
        R_ivp, M_ivp = _tov_in_r(t["K_si"], t["Gamma"], math.exp(lr))
        assert M == pytest.approx(M_ivp, rel=1e-10)
        assert R == pytest.approx(R_ivp, rel=1e-9)


def test_polytrope_newtonian_limit():
    """A light Gamma = 2 star is nearly Newtonian: R -> pi / sqrt(2 pi G / K)
    (the n = 1 Lane-Emden radius, independent of mass); general relativity makes
    it smaller, by less than its compactness 2GM/Rc^2."""
    K = 2.26e-3                            # R_newt = 7.3 km: a weak-field star
    R_newt = math.pi / math.sqrt(2.0 * math.pi * G / K)
    R, M = S.polytrope_star(K, 2.0, 0.1 * S.M_SUN)
    comp = 2.0 * G * M / (R * C_LIGHT ** 2)
    assert comp < 0.06
    assert M == pytest.approx(0.1 * S.M_SUN, rel=1e-9)
    assert -comp < R / R_newt - 1.0 < 0.0


def test_polytrope_rejects_impossible_mass():
    t = META["tov"]
    with pytest.raises(BhxError):
        S.polytrope_star(t["K_si"], t["Gamma"], 5.0 * S.M_SUN)


def test_neutron_star_scene():
    t = META["tov"]
    st = t["stars"][1]
    sc = S.default_scene()
    sc.jet_enabled, sc.ns_spin_hz = 1, 300.0
    S.neutron_star_scene(sc, st["R_m"], st["M_kg"])
    r_g = G * st["M_kg"] / C_LIGHT ** 2
    assert sc.ns_enabled == 1 and sc.disc_enabled == 0 and sc.jet_enabled == 0 and sc.spin == 0.0
    assert sc.M_kg == st["M_kg"] and sc.ns_R == st["R_m"] / r_g
    assert sc.ns_Omega == 2.0 * math.pi * 300.0 * r_g / C_LIGHT
    assert sc.cam_r == 40.0 * sc.ns_R
    assert sc.r_escape == 3.0 * sc.cam_r + 100.0
