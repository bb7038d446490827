"""set_frame / frame_scenes / probe_scene of blackhole_amd.animation: the
CLI's interpolation rules (src/main.cpp:79-86, 938-961), on the host."""
import math

import pytest

from blackhole_amd import animation as A, scene as S


def _base():
    sc = S.plain_frame(48, 32)
    sc.stars_enabled, sc.star_count = 1, 20000
    return sc


def test_set_frame_rules():
    sc = _base()
    sc.cam_theta = S.inclination(60.0)
    ss = A.frame_scenes(sc, 4, incl_to_deg=179.9, dist_to=600.0, fov_to=0.1, time_span=30.0)
    assert ss[0].cam_theta == pytest.approx(math.radians(60.0), rel=1e-15)
    assert ss[-1].cam_theta == pytest.approx(math.radians(179.5), rel=1e-15)   # clamped
    assert ss[-1].cam_r == 600.0 and ss[-1].cam_fov == 0.1 and ss[-1].observer_time == 30.0
    assert all(s.r_escape == sc.r_escape for s in ss)
    orbit = A.frame_scenes(sc, 4, time_span=40.0)
    assert [s.observer_time for s in orbit] == [0.0, 10.0, 20.0, 30.0]      # periodic: no repeat
    p = A.probe_scene(sc, time_span=40.0)
    assert (p.width, p.height, p.sqrt_spp, p.observer_time) == (64, 36, 1, 20.0)
