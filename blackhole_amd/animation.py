"""The CLI's animation loop (src/main.cpp:931-997) over the batched launch.

The reference renders a film frame by frame: each frame's configuration is
interpolated from the base one (`set_frame`, src/main.cpp:79-86: inclination,
distance, field of view, and the clock for an orbiting hot spot), the
exposure is metered once on a low-resolution probe of the middle frame and
then held for every frame (src/main.cpp:935-950), and each frame gets the
optional veiling glare and the display transform.  Here every frame is a
scene of ONE batched launch (bhx_batch: one persistent tracer launch,
replayed as a hipGraph), and the probe, the glare and the display transform
stay on the device (bhx_add_glare_dev, bhx_auto_exposure_dev,
bhx_develop_dev); only the 8-bit frames come back.

    frames, exposure, stats = animate(tr, scene, 120, incl_to_deg=90.0, glare=0.2)
"""
import ctypes as C
import math

import numpy as np

from . import native
from .abi import TM_HDR, Scene


def _copy(scene):
    s = Scene()
    C.pointer(s)[0] = scene
    return s


def set_frame(scene, u, *, incl_from_deg=None, incl_to_deg=None, dist_to=0.0, fov_to=0.0, time_span=0.0):
    """The base scene at fraction u in [0, 1] of the sequence (main.cpp:79-86):
    inclination (degrees, clamped to [0.5, 179.5]), camera distance and field
    of view interpolated linearly when an end value is given, and the observer
    clock at u * time_span.  r_escape stays the base scene's."""
    s = _copy(scene)
    if incl_to_deg is not None:
        i0 = math.degrees(scene.cam_theta) if incl_from_deg is None else incl_from_deg
        s.cam_theta = min(max(i0 + u * (incl_to_deg - i0), 0.5), 179.5) * math.pi / 180.0
    if dist_to > 0.0:
        s.cam_r = scene.cam_r + u * (dist_to - scene.cam_r)
    if fov_to > 0.0:
        s.cam_fov = scene.cam_fov + u * (fov_to - scene.cam_fov)
    s.observer_time = u * time_span
    return s


def frame_scenes(scene, frames, **kw):
    """The sequence's scenes.  A camera sweep reaches its end value on the
    last frame (u = f / (frames - 1)); a pure hot-spot orbit is periodic and
    must not repeat its first frame (u = f / frames), main.cpp:956-961."""
    sweeping = kw.get("incl_to_deg") is not None or kw.get("dist_to", 0.0) > 0.0 or kw.get("fov_to", 0.0) > 0.0
    denom = max(1, frames - 1) if sweeping else frames
    return [set_frame(scene, f / denom, **kw) for f in range(frames)]


def probe_scene(scene, **kw):
    """The exposure probe (main.cpp:938-946): a sixth of the size (at least
    64 x 36), one sample per pixel, at most 40,000 stars, the middle frame."""
    p = _copy(scene)
    p.width = max(64, scene.width // 6)
    p.height = max(36, scene.height // 6)
    p.sqrt_spp = 1
    p.star_count = min(scene.star_count, 40000)
    return set_frame(p, 0.5, **kw)


def metered_exposure(tracer, scene, glare=0.0, glare_radius=3.0, **kw):
    """The locked exposure: img::auto_exposure of the probe frame (after its
    glare), computed on the device."""
    p = probe_scene(scene, **kw)
    rays, st = tracer.trace(p)
    sh, st = tracer.shade(p, rays, stats=st, per_ray=False)
    xyz = sh["xyz"]
    if glare > 0.0:
        tracer.add_glare(xyz, glare, glare_radius)
    return tracer.auto_exposure(xyz)


def animate(tracer, scene, frames, *, incl_from_deg=None, incl_to_deg=None, dist_to=0.0, fov_to=0.0,
            time_span=0.0, exposure=0.0, glare=0.0, glare_radius=3.0, tonemap=TM_HDR, log_decades=16.0,
            graph=True):
    """Render `frames` frames of the sequence as one batch.  Returns (list of
    (H, W, 3) uint8 sRGB frames, the exposure used, per-frame Stats)."""
    from .tracer import Batch
    torch = tracer.torch
    kw = dict(incl_from_deg=incl_from_deg, incl_to_deg=incl_to_deg, dist_to=dist_to, fov_to=fov_to,
              time_span=time_span)
    if exposure <= 0.0:
        exposure = metered_exposure(tracer, scene, glare=glare, glare_radius=glare_radius, **kw)
    scenes = frame_scenes(scene, frames, **kw)
    out = []
    stream = tracer._stream(None)
    with Batch(tracer, scenes) as b:
        b.run(graph=graph)
        stats, _ = b.stats()
        n = scene.width * scene.height
        for i in range(frames):
            p = b.xyz_ptr(i)
            if glare > 0.0 and glare_radius > 0.0:
                native.check(tracer.ctx, tracer.lib.bhx_add_glare_dev(tracer.ctx, p, scene.width, scene.height,
                                                                      float(glare), float(glare_radius), stream),
                             "bhx_add_glare_dev")
            rgb = torch.empty(n * 3, dtype=torch.uint8, device=f"cuda:{tracer.device}")
            native.check(tracer.ctx, tracer.lib.bhx_develop_dev(tracer.ctx, p, n, float(exposure), int(tonemap),
                                                                float(log_decades), C.c_void_p(rgb.data_ptr()),
                                                                stream), "bhx_develop_dev")
            out.append(rgb.view(scene.height, scene.width, 3).cpu().numpy())
    return out, exposure, stats
