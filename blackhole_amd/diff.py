"""Differentiable disc images: the bhx_adjoint C ABI under torch.autograd.

    dr = DiscRender(tracer, scene, K=64)
    p = torch.tensor([spin, cam_theta, eddington, cam_r], dtype=torch.float64,
                     requires_grad=True)
    I = dr(p)                               # (H, W) bolometric intensity, W m^-2 sr^-1
    loss = ((blur(I) - blur(I_target)) ** 2).sum()
    loss.backward()                         # p.grad from the discrete adjoint

The forward is the reference's disc image (src/disc.h:94-121,
src/render.cpp:280-300: I = sigma (g T)^4 / pi on disc hits, 0 elsewhere) with
the camera fixed in everything but (spin, cam_theta, eddington, cam_r); the
backward runs the hand-written adjoint kernel (blackhole_amd/csrc/adjoint.hip).
The gradient is that of the smooth part of the image: pixels whose class
changes under an infinitesimal parameter change (disc edges, the shadow's
rim) contribute boundary terms that no pointwise derivative sees.
"""
import ctypes as C
import math

import numpy as np

from . import native
from .abi import Scene

PARAMS = ("spin", "cam_theta", "eddington", "cam_r")


def _unit_cone_jet(s):
    """s re-normalised in place for phi_flux = eps_rad = 1; its j_nu0."""
    native.check(None, native.lib().bhx_setup_conical_jet(C.byref(s), 1.0, 1.0), "bhx_setup_conical_jet")
    return s.jet_j_nu0


def with_params(scene, p):
    """Copy of scene with (spin, cam_theta, eddington, cam_r) = p; r_escape
    and the rest of the camera stay as they are.  The conical jet's emission,
    j_nu0 ~ phi_flux^2 eps_rad x (a function of spin and accretion rate), is
    carried to the new spin and rate at the scene's phi_flux^2 eps_rad (and
    kept as it is, bit for bit, when spin and rate are the scene's)."""
    s = Scene()
    C.pointer(s)[0] = scene
    s.spin, s.cam_theta, s.eddington, s.cam_r = (float(v) for v in p)
    moved = (s.spin, s.eddington) != (scene.spin, scene.eddington)
    if moved and s.jet_enabled and s.jet_model == 1 and scene.jet_j_nu0 > 0.0:
        old = Scene()
        C.pointer(old)[0] = scene
        scale = scene.jet_j_nu0 / _unit_cone_jet(old)
        s.jet_j_nu0 = _unit_cone_jet(s) * scale
    return s


class DiscRender:
    """One bhx_adjoint (checkpoint slab, scratch) for frames of one size.

    The library keeps the checkpoints of the LAST forward only; backward_raw
    differentiates that one.  Every forward_raw bumps `generation`, and the
    autograd path (calling the object) records the generation and the
    parameters of its forward, so that a backward whose checkpoints were
    overwritten by a later forward first records its own forward again (the
    forward is deterministic: the same bits, so the same gradient)."""

    def __init__(self, tracer, scene, K=64):
        self.tr = tracer
        self.lib = tracer.lib
        self.scene = scene
        self.K = int(K)
        self.generation = 0
        self.adj = C.c_void_p()
        native.check(tracer.ctx, self.lib.bhx_adjoint_create(tracer.ctx, C.byref(scene), self.K,
                                                             C.byref(self.adj)), "bhx_adjoint_create")

    def close(self):
        if self.adj:
            self.lib.bhx_adjoint_destroy(self.adj)
            self.adj = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- raw entry points --------------------------------------------------
    def forward_raw(self, p, out=None):
        """I per sample (device tensor) at parameters p; records checkpoints."""
        t = self.tr.torch
        s = with_params(self.scene, p)
        native.check(self.tr.ctx, self.lib.bhx_adjoint_set_scene(self.adj, C.byref(s)), "bhx_adjoint_set_scene")
        if out is None:
            out = t.empty(s.n_rays, dtype=t.float64, device=f"cuda:{self.tr.device}")
        native.check(self.tr.ctx, self.lib.bhx_adjoint_forward(self.adj, C.c_void_p(out.data_ptr()),
                                                               self.tr._stream(None)), "bhx_adjoint_forward")
        self.generation += 1
        return out

    def backward_raw(self, w):
        """dL/d(spin, cam_theta, eddington, cam_r) for w = dL/dI per sample,
        at the parameters of the last forward_raw."""
        w = w.contiguous()
        g = (C.c_double * 4)()
        native.check(self.tr.ctx, self.lib.bhx_adjoint_backward(self.adj, C.c_void_p(w.data_ptr()), g,
                                                                self.tr._stream(None)), "bhx_adjoint_backward")
        return np.array(g[:])

    def rays(self):
        """Class, hit state (8, n) and accepted steps of the last forward (host copies)."""
        n = self.scene.n_rays
        out = {"cls": np.zeros(n, np.uint8), "y": np.zeros((8, n)), "steps": np.zeros(n, np.int32)}
        native.check(self.tr.ctx, self.lib.bhx_adjoint_copy_rays(
            self.adj, *[out[k].ctypes.data_as(C.c_void_p) for k in ("cls", "y", "steps")],
            self.tr._stream(None)), "bhx_adjoint_copy_rays")
        return out

    def autotune(self, w, Ks=(1, 16, 32, 64, 128, 256), reps=2):
        """Pick the fastest checkpoint interval among Ks for this frame (timed
        gradient passes with dL/dI = w), then the recording forward's waves per
        SIMD at that K; returns (K, ms per pass).  The next backward needs a
        new forward."""
        w = w.contiguous()
        arr = (C.c_int32 * len(Ks))(*Ks)
        k, ms = C.c_int32(), C.c_double()
        native.check(self.tr.ctx, self.lib.bhx_adjoint_autotune(self.adj, C.c_void_p(w.data_ptr()), arr, len(Ks),
                                                                int(reps), C.byref(k), C.byref(ms),
                                                                self.tr._stream(None)), "bhx_adjoint_autotune")
        self.K = k.value
        self.generation += 1
        return k.value, ms.value

    def set_launch(self, q_chunk=0, refill=0):
        """The recording forward's queue chunk and refill threshold (0 keeps a value)."""
        native.check(self.tr.ctx, self.lib.bhx_adjoint_set_launch(self.adj, int(q_chunk), int(refill)),
                     "bhx_adjoint_set_launch")

    def set_forward_waves(self, waves=0):
        """Waves per SIMD of the recording forward (0: as many as fit)."""
        native.check(self.tr.ctx, self.lib.bhx_adjoint_set_forward_waves(self.adj, int(waves)),
                     "bhx_adjoint_set_forward_waves")

    def info(self):
        ck, disc, mism = C.c_int64(), C.c_int64(), C.c_int64()
        f, b = C.c_double(), C.c_double()
        wv = C.c_int32()
        native.check(self.tr.ctx, self.lib.bhx_adjoint_info(self.adj, C.byref(ck), C.byref(disc), C.byref(mism),
                                                            C.byref(f), C.byref(b)), "bhx_adjoint_info")
        native.check(self.tr.ctx, self.lib.bhx_adjoint_get_forward_waves(self.adj, C.byref(wv)),
                     "bhx_adjoint_get_forward_waves")
        return {"ck_bytes": ck.value, "lit_rays": disc.value, "mismatches": mism.value,
                "fwd_ms": f.value, "bwd_ms": b.value, "fwd_waves": wv.value}

    # ---- autograd ------------------------------------------------------------
    def __call__(self, p):
        return _DiscImage.apply(p, self)


class DiscBatch:
    """One bhx_adjoint over a batch of frames of one size and one physics that
    differ in their camera only (a camera sweep): one forward traces every
    frame's rays in one persistent launch, one backward gives each frame its
    own gradient.

        db = DiscBatch(tracer, frames, K=64)
        I = db.forward_raw()          # (n_frames, samples) device tensor
        g = db.backward_raw(w)        # (n_frames, 4): dL_f/d(spin, cam_theta, eddington, cam_r)

    w holds dL/dI for every frame's samples (same layout as I); frame f's row
    of g is the gradient of the part of L its own samples carry."""

    def __init__(self, tracer, scenes, K=64):
        self.tr = tracer
        self.lib = tracer.lib
        self.scenes = list(scenes)
        self.n = len(self.scenes)
        self.n_per = self.scenes[0].n_rays
        self.K = int(K)
        self.generation = 0
        self.adj = C.c_void_p()
        arr = (Scene * self.n)(*self.scenes)
        native.check(tracer.ctx, self.lib.bhx_adjoint_create_batch(tracer.ctx, arr, self.n, self.K,
                                                                   C.byref(self.adj)), "bhx_adjoint_create_batch")

    def close(self):
        if self.adj:
            self.lib.bhx_adjoint_destroy(self.adj)
            self.adj = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward_raw(self, out=None):
        """I of every frame's samples, (n_frames, samples) on the device; records checkpoints."""
        t = self.tr.torch
        if out is None:
            out = t.empty((self.n, self.n_per), dtype=t.float64, device=f"cuda:{self.tr.device}")
        native.check(self.tr.ctx, self.lib.bhx_adjoint_forward(self.adj, C.c_void_p(out.data_ptr()),
                                                               self.tr._stream(None)), "bhx_adjoint_forward")
        self.generation += 1
        return out

    def backward_raw(self, w):
        """Per-frame dL/d(spin, cam_theta, eddington, cam_r), (n_frames, 4),
        for w = dL/dI of the last forward's samples."""
        w = w.contiguous()
        g = (C.c_double * (4 * self.n))()
        native.check(self.tr.ctx, self.lib.bhx_adjoint_backward_batch(self.adj, C.c_void_p(w.data_ptr()), g,
                                                                      self.tr._stream(None)),
                     "bhx_adjoint_backward_batch")
        return np.array(g[:]).reshape(self.n, 4)

    def rays(self):
        """Class and accepted steps of the last forward's rays, (n_frames, samples) host copies."""
        out = {"cls": np.zeros((self.n, self.n_per), np.uint8), "steps": np.zeros((self.n, self.n_per), np.int32)}
        native.check(self.tr.ctx, self.lib.bhx_adjoint_copy_rays(
            self.adj, out["cls"].ctypes.data_as(C.c_void_p), None, out["steps"].ctypes.data_as(C.c_void_p),
            self.tr._stream(None)), "bhx_adjoint_copy_rays")
        return out

    autotune = DiscRender.autotune
    set_launch = DiscRender.set_launch
    info = DiscRender.info


def _image_from_samples(tr, I, scene):
    S2 = max(1, scene.sqrt_spp) ** 2
    return I.view(scene.height, scene.width, S2).mean(-1)


import torch  # noqa: E402  (torch is the autograd host; the arithmetic is in libbhx.so)


class _DiscImage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, dr):
        params = p.detach().cpu().numpy().copy()
        I = dr.forward_raw(params)
        ctx.dr = dr
        ctx.device = p.device
        ctx.params = params
        ctx.generation = dr.generation
        return _image_from_samples(dr.tr, I, dr.scene)

    @staticmethod
    def backward(ctx, gI):
        dr = ctx.dr
        sc = dr.scene
        if dr.generation != ctx.generation:
            # a later forward replaced this one's checkpoints: record it again
            dr.forward_raw(ctx.params)
        S2 = max(1, sc.sqrt_spp) ** 2
        w = (gI.to(torch.float64).unsqueeze(-1).expand(sc.height, sc.width, S2) / S2).contiguous().view(-1)
        g = dr.backward_raw(w)
        return torch.tensor(g, dtype=torch.float64, device=ctx.device), None


def gaussian_blur(img, sigma=2.0):
    """Separable Gaussian of radius ceil(3 sigma), clamp-to-edge, normalised
    weights - the blur of img::add_glare (src/image.cpp:61-93) - over the
    last two axes of a (..., H, W) tensor, differentiable."""
    r = max(1, int(math.ceil(3.0 * sigma)))
    i = torch.arange(-r, r + 1, dtype=torch.float64)
    k = torch.exp(-0.5 * (i ** 2) / (sigma * sigma))
    k = (k / k.sum()).tolist()
    H, W = img.shape[-2:]
    # clamp-to-edge: the border rows/columns repeated r times; each pass a
    # sum of shifted slices (no gathers, so no scattered adds backward)
    xp = torch.cat([img[..., :1].expand(*img.shape[:-1], r), img, img[..., -1:].expand(*img.shape[:-1], r)], -1)
    tmp = k[0] * xp[..., 0:W]
    for t in range(1, 2 * r + 1):
        tmp = tmp + k[t] * xp[..., t:t + W]
    lead = tmp.shape[:-2]
    tp = torch.cat([tmp[..., :1, :].expand(*lead, r, W), tmp, tmp[..., -1:, :].expand(*lead, r, W)], -2)
    out = k[0] * tp[..., 0:H, :]
    for t in range(1, 2 * r + 1):
        out = out + k[t] * tp[..., t:t + H, :]
    return out


def gaussian_blur_np(img, sigma=2.0):
    """gaussian_blur on a numpy array (the same sums, for CPU checks)."""
    r = max(1, int(math.ceil(3.0 * sigma)))
    i = np.arange(-r, r + 1)
    k = np.exp(-0.5 * (i.astype(np.float64) ** 2) / (sigma * sigma))
    k = k / k.sum()
    H, W = img.shape
    ix = np.clip(np.arange(W)[:, None] + i[None, :], 0, W - 1)
    tmp = (img[:, ix] * k).sum(-1)
    iy = np.clip(np.arange(H)[:, None] + i[None, :], 0, H - 1)
    return (tmp[iy, :] * k[None, :, None]).sum(1)
