"""The CLI's animation loop (src/main.cpp:931-997) as one batched launch with
the display pipeline on the device (blackhole_amd.animation), against the
same loop run by the oracle: frame scenes from set_frame, the exposure
metered on the probe of the middle frame (with its glare) and held, each
frame's glare and display transform.  Needs an MI355X."""
import math

import numpy as np
import pytest

import oracle
from blackhole_amd import animation as A, scene as S

pytestmark = pytest.mark.gpu


def _base():
    sc = S.plain_frame(48, 32)
    sc.stars_enabled, sc.star_count = 1, 20000
    sc.hs_enabled, sc.hs_omega = 1, 1 / (8 * math.sqrt(8) + 0.9)   # hot spot near r = 8
    return sc


def _oracle_film(sc, frames, glare, radius, **kw):
    p = A.probe_scene(sc, **kw)
    xyz, _ = oracle.render(p)
    e = oracle.auto_exposure(oracle.add_glare(xyz, glare, radius) if glare > 0 else xyz)
    out = []
    for f in A.frame_scenes(sc, frames, **kw):
        xyz, _ = oracle.render(f)
        if glare > 0:
            xyz = oracle.add_glare(xyz, glare, radius)
        out.append(oracle.develop(xyz, e, 4, 16.0).reshape(f.height, f.width, 3))
    return out, e


@pytest.mark.parametrize("kw", [
    dict(incl_to_deg=80.0, time_span=200.0),           # camera sweep + hot-spot clock
    dict(time_span=300.0),                              # periodic orbit: u = f / frames
])
def test_film_matches_oracle(tracer, kw):
    sc = _base()
    frames, e, stats = A.animate(tracer, sc, 5, glare=0.2, glare_radius=2.0, **kw)
    ref, e_ref = _oracle_film(sc, 5, 0.2, 2.0, **kw)
    print(f"[film {kw}] exposure {e} oracle {e_ref}", flush=True)
    assert abs(e - e_ref) <= 1e-9 * abs(e_ref)
    for i, (a, b) in enumerate(zip(frames, ref)):
        d = int(np.abs(a.astype(int) - b.astype(int)).max())
        assert d <= 1, (i, d)
    assert len(stats) == 5 and all(st.rays == sc.width * sc.height for st in stats)
    # the frames differ (the camera or the hot spot moves)
    assert not np.array_equal(frames[0], frames[-1])
