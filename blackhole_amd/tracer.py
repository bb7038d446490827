"""Host-side mirror of the reference's frame API over the HIP tracer.

`Tracer.render(scene)` is the Python face of bh::render (src/render.h:72):
it returns the absolute CIE XYZ frame and the accumulated RenderStats.  The
stage calls (trace / shade / auto_exposure / develop) keep everything in
device memory; torch provides the allocations and the stream, the arithmetic
is all in libbhx.so.
"""
import ctypes as C

import numpy as np

from . import native
from .abi import TM_HDR, Bytes, Launch, Rays, Scene, Shaded, Stats


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class Tracer:
    """One bhx_context on one GPU."""

    def __init__(self, device=0):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise native.BhxError("no GPU visible: the HIP tracer has no CPU path")
        self.device = device
        self.lib = native.lib()
        self.ctx = C.c_void_p()
        torch.cuda.set_device(device)
        native.check(None, self.lib.bhx_create(device, C.byref(self.ctx)), "bhx_create")

    def close(self):
        if self.ctx:
            self.lib.bhx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        return C.c_void_p(s.cuda_stream)

    # ---- configuration -----------------------------------------------------
    def set_catalogue(self, x=None, y=None, z=None, T=None, irr=None):
        """Load an explicit star catalogue (host arrays); no arguments = the
        reference generator StarField(star_count, seed)."""
        if x is None:
            native.check(self.ctx, self.lib.bhx_set_catalogue(self.ctx, 0, None, None, None, None, None),
                         "bhx_set_catalogue")
            return
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, T, irr)]
        native.check(self.ctx, self.lib.bhx_set_catalogue(
            self.ctx, len(arrs[0]), *[a.ctypes.data_as(C.c_void_p) for a in arrs]), "bhx_set_catalogue")

    def set_launch(self, block=0, waves_per_simd=0, chunk=0, refill=0, queue_order=0):
        """Launch shape of single-frame traces (0 keeps a field); queue_order is
        abi.ORDER_CENTRE or abi.ORDER_COST (the default)."""
        native.check(self.ctx, self.lib.bhx_set_launch(
            self.ctx, C.byref(Launch(block, waves_per_simd, chunk, refill, queue_order))), "bhx_set_launch")

    def get_launch(self):
        l = Launch()
        native.check(self.ctx, self.lib.bhx_get_launch(self.ctx, C.byref(l)), "bhx_get_launch")
        return {"block": l.block, "waves_per_simd": l.waves_per_simd, "chunk": l.chunk,
                "refill": l.refill, "queue_order": l.queue_order}

    # ---- drop-in frame API ---------------------------------------------------
    def render(self, scene):
        """bh::render: (xyz[H, W, 3] float64 numpy, Stats)."""
        xyz = np.zeros((scene.height, scene.width, 3))
        st = Stats()
        native.check(self.ctx, self.lib.bhx_render(self.ctx, C.byref(scene),
                                                   xyz.ctypes.data_as(C.c_void_p), C.byref(st)),
                     "bhx_render")
        return xyz, st

    # ---- device stages -------------------------------------------------------
    def alloc_rays(self, scene, full=True):
        t = self.torch
        n = scene.n_rays
        dev = f"cuda:{self.device}"
        out = {"cls": t.empty(n, dtype=t.uint8, device=dev),
               "y": t.empty((8, n), dtype=t.float64, device=dev),
               "jet": t.empty(n, dtype=t.float64, device=dev)}
        if full:
            out.update(steps=t.empty(n, dtype=t.int32, device=dev),
                       attempts=t.empty(n, dtype=t.int32, device=dev),
                       norm_err=t.empty(n, dtype=t.float64, device=dev),
                       carter=t.empty(n, dtype=t.float64, device=dev))
        return out

    @staticmethod
    def _rays_struct(r):
        return Rays(*[(r[k].data_ptr() if r.get(k) is not None else None)
                      for k in ("cls", "y", "jet", "steps", "attempts", "norm_err", "carter")])

    def trace(self, scene, rays=None, stats=None, stream=None):
        """trace() of every sample into device arrays; returns (rays, stats)."""
        if rays is None:
            rays = self.alloc_rays(scene)
        st = stats if stats is not None else Stats()
        native.check(self.ctx, self.lib.bhx_trace_dev(self.ctx, C.byref(scene),
                                                      C.byref(self._rays_struct(rays)), C.byref(st),
                                                      self._stream(stream)), "bhx_trace_dev")
        return rays, st

    def trace_async(self, scene, rays, stream=None):
        """Stream-ordered trace without a host synchronisation (for timing loops)."""
        native.check(self.ctx, self.lib.bhx_trace_dev(self.ctx, C.byref(scene),
                                                      C.byref(self._rays_struct(rays)), None,
                                                      self._stream(stream)), "bhx_trace_dev")

    def shade(self, scene, rays, out=None, stats=None, stream=None, per_ray=True):
        t = self.torch
        dev = f"cuda:{self.device}"
        if out is None:
            out = {"xyz": t.empty((scene.height, scene.width, 3), dtype=t.float64, device=dev)}
            if per_ray:
                n = scene.n_rays
                out.update(ray_xyz=t.empty((3, n), dtype=t.float64, device=dev),
                           g=t.empty(n, dtype=t.float64, device=dev),
                           dir=t.empty((3, n), dtype=t.float64, device=dev))
        st = stats if stats is not None else Stats()
        sh = Shaded(*[(out[k].data_ptr() if out.get(k) is not None else None)
                      for k in ("xyz", "ray_xyz", "g", "dir")])
        native.check(self.ctx, self.lib.bhx_shade_dev(self.ctx, C.byref(scene),
                                                      C.byref(self._rays_struct(rays)), C.byref(sh),
                                                      C.byref(st), self._stream(stream)), "bhx_shade_dev")
        return out, st

    def shade_async(self, scene, rays, out, stream=None):
        sh = Shaded(*[(out[k].data_ptr() if out.get(k) is not None else None)
                      for k in ("xyz", "ray_xyz", "g", "dir")])
        native.check(self.ctx, self.lib.bhx_shade_dev(self.ctx, C.byref(scene),
                                                      C.byref(self._rays_struct(rays)), C.byref(sh),
                                                      None, self._stream(stream)), "bhx_shade_dev")

    def auto_exposure(self, xyz, key=1.0, stream=None):
        e = C.c_double()
        n = xyz.numel() // 3
        native.check(self.ctx, self.lib.bhx_auto_exposure_dev(self.ctx, _ptr(xyz), n, key, C.byref(e),
                                                              self._stream(stream)), "bhx_auto_exposure_dev")
        return e.value

    def add_glare(self, xyz, strength, radius_px=3.0, stream=None):
        """img::add_glare (src/image.cpp:47-106) in place on a contiguous
        (H, W, 3) float64 device tensor; returns it."""
        if xyz.dim() != 3 or xyz.shape[2] != 3 or not xyz.is_contiguous():
            raise ValueError("add_glare needs a contiguous (H, W, 3) tensor")
        H, W = int(xyz.shape[0]), int(xyz.shape[1])
        native.check(self.ctx, self.lib.bhx_add_glare_dev(self.ctx, _ptr(xyz), W, H, float(strength),
                                                          float(radius_px), self._stream(stream)),
                     "bhx_add_glare_dev")
        return xyz

    def develop(self, xyz, exposure, tonemap=TM_HDR, log_decades=16.0, stream=None):
        t = self.torch
        n = xyz.numel() // 3
        rgb = t.empty(n * 3, dtype=t.uint8, device=xyz.device)
        native.check(self.ctx, self.lib.bhx_develop_dev(self.ctx, _ptr(xyz), n, exposure, tonemap,
                                                        log_decades, _ptr(rgb), self._stream(stream)),
                     "bhx_develop_dev")
        return rgb.view(-1, 3)

    def develop_auto(self, xyz, key=1.0, tonemap=TM_HDR, log_decades=16.0, exposure_out=None, stream=None):
        """auto_exposure then develop (src/image.cpp:22-45, 119-191) with the
        exposure kept on the device: the same sRGB8 as
        develop(xyz, auto_exposure(xyz, key), ...), queued on the stream with
        no host synchronisation.  exposure_out: an optional one-element
        float64 device tensor that receives the exposure."""
        t = self.torch
        n = xyz.numel() // 3
        rgb = t.empty(n * 3, dtype=t.uint8, device=xyz.device)
        e_ptr = None
        if exposure_out is not None:
            if exposure_out.dtype != t.float64 or exposure_out.numel() < 1 or not exposure_out.is_cuda:
                raise ValueError("exposure_out must be a float64 device tensor")
            e_ptr = _ptr(exposure_out)
        native.check(self.ctx, self.lib.bhx_develop_auto_dev(self.ctx, _ptr(xyz), n, key, tonemap, log_decades,
                                                             _ptr(rgb), e_ptr, self._stream(stream)),
                     "bhx_develop_auto_dev")
        return rgb.view(-1, 3)

    # ---- PNG / APNG (img::write_png / write_apng, src/image.cpp:193-778) -----
    def _take(self, out):
        try:
            return C.string_at(out.data, out.size)
        finally:
            self.lib.bhx_bytes_free(C.byref(out))

    def png_encode(self, rgb, stream=None):
        """The reference writer's PNG bytes of an (H, W, 3) uint8 image: a
        numpy array (host) or a contiguous device tensor."""
        out = Bytes()
        if isinstance(rgb, np.ndarray):
            a = np.ascontiguousarray(rgb, dtype=np.uint8)
            h, w = a.shape[:2]
            rc = self.lib.bhx_png_encode(self.ctx, a.ctypes.data_as(C.c_void_p), w, h, C.byref(out))
        else:
            if rgb.dtype != self.torch.uint8 or rgb.dim() != 3 or not rgb.is_contiguous():
                raise ValueError("png_encode needs a contiguous (H, W, 3) uint8 tensor")
            h, w = int(rgb.shape[0]), int(rgb.shape[1])
            rc = self.lib.bhx_png_encode_dev(self.ctx, _ptr(rgb), w, h, self._stream(stream), C.byref(out))
        native.check(self.ctx, rc, "bhx_png_encode")
        return self._take(out)

    def apng_encode(self, frames, fps, loops=0, stream=None):
        """The reference writer's APNG bytes of a sequence of (H, W, 3) uint8
        frames: numpy arrays (host) or contiguous device tensors."""
        out = Bytes()
        n = len(frames)
        if n == 0:
            raise ValueError("apng_encode needs at least one frame")
        h, w = int(frames[0].shape[0]), int(frames[0].shape[1])
        if isinstance(frames[0], np.ndarray):
            keep = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames]
            ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in keep])
            rc = self.lib.bhx_apng_encode(self.ctx, ptrs, n, w, h, fps, loops, C.byref(out))
        else:
            ptrs = (C.c_void_p * n)(*[_ptr(f) for f in frames])
            rc = self.lib.bhx_apng_encode_dev(self.ctx, ptrs, n, w, h, fps, loops, self._stream(stream),
                                              C.byref(out))
        native.check(self.ctx, rc, "bhx_apng_encode")
        return self._take(out)

    def blackbody_xyz(self, T, stream=None):
        t = self.torch
        out = t.empty((T.numel(), 3), dtype=t.float64, device=T.device)
        native.check(self.ctx, self.lib.bhx_blackbody_xyz_dev(self.ctx, _ptr(T), T.numel(), _ptr(out),
                                                              self._stream(stream)), "bhx_blackbody_xyz_dev")
        return out


class Batch:
    """Many scenes traced as one job (bhx_batch): one persistent tracer
    launch over every scene's rays, replayed as a hipGraph."""

    def __init__(self, tracer, scenes):
        self.tr = tracer
        self.lib = tracer.lib
        self.scenes = list(scenes)
        arr = (Scene * len(self.scenes))(*self.scenes)
        self.b = C.c_void_p()
        native.check(tracer.ctx, self.lib.bhx_batch_create(tracer.ctx, arr, len(self.scenes), C.byref(self.b)),
                     "bhx_batch_create")

    def close(self):
        if self.b:
            self.lib.bhx_batch_destroy(self.b)
            self.b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, graph=True, stream=None):
        native.check(self.tr.ctx, self.lib.bhx_batch_run(self.b, int(bool(graph)), self.tr._stream(stream)),
                     "bhx_batch_run")

    def stats(self, stream=None):
        """Per-scene Stats of the last run and the batch's device time (ms)."""
        st = (Stats * len(self.scenes))()
        ms = C.c_double()
        native.check(self.tr.ctx, self.lib.bhx_batch_stats(self.b, st, C.byref(ms), self.tr._stream(stream)),
                     "bhx_batch_stats")
        return list(st), ms.value

    def xyz_ptr(self, i):
        """Device address of scene i's (H, W, 3) float64 XYZ pixels after a run
        (valid until the next run or close)."""
        p = C.c_void_p()
        native.check(self.tr.ctx, self.lib.bhx_batch_scene(self.b, i, None, None, None, C.byref(p)),
                     "bhx_batch_scene")
        return p

    def scene(self, i, stream=None):
        """Host copies of scene i's results: cls, y (8, n), jet, steps, attempts, xyz (H, W, 3)."""
        sc = self.scenes[i]
        n = sc.n_rays
        out = {"cls": np.zeros(n, np.uint8), "y": np.zeros((8, n)), "jet": np.zeros(n),
               "steps": np.zeros(n, np.int32), "attempts": np.zeros(n, np.int32),
               "xyz": np.zeros((sc.height, sc.width, 3))}
        native.check(self.tr.ctx, self.lib.bhx_batch_copy_scene(
            self.b, i, *[out[k].ctypes.data_as(C.c_void_p) for k in ("cls", "y", "jet", "steps", "attempts", "xyz")],
            self.tr._stream(stream)), "bhx_batch_copy_scene")
        return out

    def set_launch(self, block=0, waves_per_simd=0, chunk=0, refill=0):
        native.check(self.tr.ctx, self.lib.bhx_batch_set_launch(
            self.b, C.byref(Launch(block, waves_per_simd, chunk, refill))), "bhx_batch_set_launch")

    def get_launch(self):
        l = Launch()
        native.check(self.tr.ctx, self.lib.bhx_batch_get_launch(self.b, C.byref(l)), "bhx_batch_get_launch")
        return {"block": l.block, "waves_per_simd": l.waves_per_simd, "chunk": l.chunk, "refill": l.refill}

    def autotune(self, reps=2, stream=None):
        """Pick the fastest launch shape for this batch; returns (shape, ms)."""
        l = Launch()
        ms = C.c_double()
        native.check(self.tr.ctx, self.lib.bhx_batch_autotune(self.b, reps, C.byref(l), C.byref(ms),
                                                              self.tr._stream(stream)), "bhx_batch_autotune")
        return {"block": l.block, "waves_per_simd": l.waves_per_simd, "chunk": l.chunk,
                "refill": l.refill}, ms.value
