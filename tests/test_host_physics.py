"""Host scene preparation in the C-ABI library (jets, TOV, neutron-star scene)
against the values the reference build wrote into the golden fixtures, plus
the closed forms that pin the TOV solver where no fixture reaches.  CPU only:
these entry points make no GPU call."""
import math

import pytest

from blackhole_amd import scene as S
from blackhole_amd.native import BhxError
from test_oracle_golden import META

G, C_LIGHT = 6.67430e-11, 2.99792458e8


@pytest.mark.parametrize("case", META["jets"], ids=lambda c: f"a{c['spin']}_edd{c['eddington']}")
def test_setup_jet_matches_reference(case):
    sc = S.default_scene()
    sc.M_kg, sc.spin, sc.eddington = 1e9 * S.M_SUN, case["spin"], case["eddington"]
    sc.disc_r_out, sc.jet_z_top, sc.jet_enabled = 20.0, case["z_top"], 1
    S.setup_jet(sc)
    assert sc.jet_enabled == 1
    assert sc.jet_a == case["spin"]
    for k in ("j_nu0", "r_bound", "Omega_F", "z_sat"):
        assert getattr(sc, "jet_" + k) == case[k], k


def test_setup_jet_golden_scene_fields():
    """The quasar scene's stored jet fields came from the reference's setup."""
    d = META["scenes"]["quasar_jets"]["scene"]
    sc = S.default_scene()
    for k in ("M_kg", "spin", "eddington", "disc_r_out", "disc_sense", "jet_z_top", "jet_gamma_max"):
        setattr(sc, k, d[k])
    sc.jet_enabled = 1
    S.setup_jet(sc)
    for k in ("jet_z_sat", "jet_a", "jet_Omega_F", "jet_j_nu0", "jet_r_bound"):
        assert getattr(sc, k) == d[k], k


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
        assert R == st["R_m"] and M == st["M_kg"]


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
