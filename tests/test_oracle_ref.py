"""The oracle against the reference compiled from its own sources (oracle/_ref).

Runs wherever oracle/_ref/libbhref.so exists (built by oracle/Makefile from
/root/reference/src).  Equality is bitwise: same algorithm, same operation
order, same glibc, no FMA contraction in either build.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from blackhole_amd import scene as S

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="reference shim not built")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def check_scene(sc, stride=3):
    xo, so = oracle.render(sc)
    xr, sr = oracle.ref_render(sc)
    assert np.array_equal(bits(xo), bits(xr))
    for k in ("rays", "steps", "captured", "disc_hits", "sun_hits", "star_hits", "jet_rays",
              "escaped", "exhausted", "max_norm_error", "max_carter_drift"):
        assert getattr(so, k) == getattr(sr, k), k
    ids = np.arange(0, sc.n_rays, stride)
    a = oracle.trace_list(sc, ids)
    b = oracle.ref_trace_list(sc, ids)
    for k in ("cls", "y", "jet", "steps", "ray_xyz"):
        assert np.array_equal(bits(a[k]), bits(b[k])), k
    disc = a["cls"] == 1
    assert np.array_equal(bits(a["g"][disc]), bits(b["g"][disc]))


def test_plain_frame_small():
    check_scene(S.plain_frame(40, 40))


def test_starry_frame_small_with_reference_stars():
    check_scene(S.starry_frame(40, 40))


def test_spp_and_hotspot():
    sc = S.plain_frame(24, 16)
    sc.sqrt_spp = 2
    sc.hs_enabled, sc.hs_omega, sc.observer_time = 1, 1 / (8 * math.sqrt(8) + 0.9), 40.0
    check_scene(sc, 1)


def test_jets_quasar_preset():
    sc = S.default_scene()
    sc.M_kg, sc.spin, sc.cam_r, sc.disc_r_out, sc.eddington = 1e9 * S.M_SUN, 0.998, 80.0, 20.0, 0.5
    sc.cam_theta, sc.cam_fov = S.inclination(84), 0.9
    sc.jet_enabled, sc.jet_z_top, sc.jet_gamma_max = 1, 130.0, 10.0
    sc.width, sc.height, sc.sqrt_spp, sc.star_count = 32, 18, 1, 20000
    oracle.ref().bhref_setup_jet(C.byref(sc))
    sc.r_escape = S.escape_radius(sc)
    check_scene(sc, 1)


def test_neutron_star_with_caps():
    sc = S.default_scene()
    sc.M_kg, sc.spin, sc.disc_enabled, sc.ns_enabled = 1.4 * S.M_SUN, 0.0, 0, 1
    sc.ns_R, sc.ns_caps, sc.ns_Omega = 6.9, 1, 0.002
    sc.cam_r = 40 * sc.ns_R
    sc.cam_fov = 5 * math.atan2(sc.ns_R / math.sqrt(1 - 2 / sc.ns_R), sc.cam_r)
    sc.cam_theta = S.inclination(60)
    sc.width, sc.height, sc.sqrt_spp, sc.star_count = 32, 32, 1, 20000
    sc.r_escape = S.escape_radius(sc)
    check_scene(sc, 1)


def test_sun():
    sc = S.default_scene()
    sc.M_kg, sc.spin, sc.cam_r, sc.disc_r_out, sc.eddington = 4.297e6 * S.M_SUN, 0.9, 150.0, 26.0, 1e-6
    sc.cam_theta, sc.cam_fov, sc.cam_yaw = S.inclination(82), 0.55, 0.2
    sc.width, sc.height, sc.sqrt_spp, sc.star_count = 32, 18, 1, 20000
    sc.sun_enabled, sc.sun_radius = 1, 30.0
    sc.sun_centre[1] = 60.0
    sc.r_escape = S.escape_radius(sc)
    check_scene(sc, 1)


def test_display_transform():
    rng = np.random.default_rng(5)
    xyz = np.abs(rng.lognormal(0, 3, size=(5000, 3))) * 1e-3
    xyz[::97] = 0.0
    xyz[5, 1] = np.nan
    e_o = oracle.auto_exposure(xyz)
    e_r = oracle.ref().bhref_auto_exposure(xyz.ctypes.data_as(C.c_void_p), len(xyz), 1.0)
    assert e_o == e_r
    for tm in range(5):
        a = oracle.develop(xyz, e_o, tm, 16.0)
        b = np.zeros_like(a)
        oracle.ref().bhref_develop(xyz.ctypes.data_as(C.c_void_p), len(xyz), e_o, tm, 16.0,
                                   b.ctypes.data_as(C.c_void_p))
        assert np.array_equal(a, b), tm
