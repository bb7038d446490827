"""Degenerate frame shapes and invalid scenes (needs an MI355X).

Frames of 1x1, 1x7, 7x1, 13x5 at 4 samples per pixel and 67x3: 8x8 tiles cut
by the frame edge, ray counts that are not a multiple of the 64-lane wave and
far below one persistent grid.  Each is checked ray by ray against the oracle
with the bar of tests/parity.py and pixel by pixel (sRGB within 1 LSB under
the oracle's exposure).  A frame without pixels is refused with an error
before any launch; tolerances are taken as given, as the reference does.
"""
import math

import numpy as np
import pytest

import oracle
from blackhole_amd import scene as S
from blackhole_amd.native import BhxError

from parity import compare, gpu_rays_numpy, srgb_lsb_diff

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1), (1, 7, 1), (7, 1, 1), (13, 5, 2), (67, 3, 1)]


def _scene(w, h, sqrt_spp):
    # a wide field so that even one pixel's rays cross disc, shadow and sky
    sc = S.plain_scene(spin=0.9, incl_deg=80.0, r_cam=500.0, fov=2.0 * math.atan(20.0 / 500.0), width=w,
                       height=h, eddington=0.1, disc_r_out=20.0, sqrt_spp=sqrt_spp)
    return sc


@pytest.mark.parametrize("w,h,s", SHAPES, ids=[f"{w}x{h}_spp{s * s}" for w, h, s in SHAPES])
def test_small_frame(tracer, w, h, s):
    sc = _scene(w, h, s)
    print(f"[small {w}x{h} spp {s * s}] n_rays={sc.n_rays}", flush=True)
    rays, st = tracer.trace(sc)
    sh, st = tracer.shade(sc, rays, stats=st)
    g = gpu_rays_numpy(rays, sh)
    ids = np.arange(sc.n_rays)
    c = oracle.trace_list(sc, ids)
    rep = compare(sc, ids, g, c)
    print(f"[small {w}x{h}] " + " ".join(f"{k}={v}" for k, v in rep.items()), flush=True)
    assert rep["fail"] == 0, rep
    assert st.rays == sc.n_rays
    assert np.bincount(g["cls"], minlength=6).sum() == sc.n_rays
    xyz_g = sh["xyz"].cpu().numpy()
    xyz_c, _ = oracle.render(sc)
    assert xyz_g.shape == xyz_c.shape == (h, w, 3)
    if xyz_c[..., 1].max() > 0:
        e_c = oracle.auto_exposure(xyz_c)
        assert tracer.auto_exposure(sh["xyz"]) == pytest.approx(e_c, rel=1e-9)
        assert srgb_lsb_diff(xyz_g, xyz_c, e_c) <= 1


@pytest.mark.parametrize("field,value", [("width", 0), ("height", -3)])
def test_empty_frame_is_refused(tracer, field, value):
    sc = _scene(8, 8, 1)
    rays = tracer.alloc_rays(sc)   # valid buffers: the library itself must refuse the scene
    setattr(sc, field, value)
    with pytest.raises(BhxError, match="must be > 0"):
        tracer.trace(sc, rays=rays)
    # the context stays usable after the refusal
    ok = _scene(4, 4, 1)
    rays, st = tracer.trace(ok)
    assert st.rays == ok.n_rays


@pytest.mark.parametrize("rtol,atol", [(0.0, 1e-8), (0.0, 0.0)], ids=["absolute_only", "both_zero"])
def test_tolerances_taken_as_given(tracer, rtol, atol):
    """The reference's Stepper takes any tolerances (render.cpp:80): rtol = 0 is a
    pure absolute tolerance; with both 0 every error is 0/0, every attempt is
    rejected and each ray ends as captured after 40 retries (render.cpp:135)."""
    sc = _scene(13, 5, 1)
    sc.rtol, sc.atol = rtol, atol
    rays, st = tracer.trace(sc)
    sh, st = tracer.shade(sc, rays, stats=st)
    g = gpu_rays_numpy(rays, sh)
    ids = np.arange(sc.n_rays)
    c = oracle.trace_list(sc, ids)
    rep = compare(sc, ids, g, c)
    print(f"[tol {rtol} {atol}] " + " ".join(f"{k}={v}" for k, v in rep.items()), flush=True)
    assert rep["fail"] == 0, rep
    assert np.array_equal(g["steps"], c["steps"]) or rtol == 0.0 and atol > 0.0
    if atol == 0.0:
        assert (c["cls"] == 0).all() and (g["cls"] == 0).all()
        assert (g["steps"] == 0).all()
