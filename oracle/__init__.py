"""CPU oracle of the per-pixel Kerr tracer -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this package, and only as the checker or the timed CPU baseline; the product
(blackhole_amd/) never imports it.

Two libraries are wrapped:
  lib/liboracle.so  the repository's own CPU restatement (src/oracle.cpp) of the
                    reference's algorithm, bitwise-pinned to the reference;
  _ref/libbhref.so  the reference itself, compiled from its sources by
                    oracle/Makefile where /root/reference exists (optional).

Parity status per function: every function of src/oracle.cpp is pinned, see
the header of src/oracle.cpp and tests/test_oracle_ref.py / test_oracle_golden.py.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

from blackhole_amd.abi import Scene, Stats

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "liboracle.so")
REF = os.path.join(HERE, "_ref", "libbhref.so")

_lib = None
_ref = None

P = C.c_void_p
D = C.c_double
DP = C.POINTER(C.c_double)


class Catalogue(C.Structure):
    _fields_ = [("n", C.c_int32), ("x", P), ("y", P), ("z", P), ("temperature", P),
                ("irradiance", P)]


class OracleRays(C.Structure):
    _fields_ = [("cls", P), ("y", P), ("jet", P), ("steps", P), ("attempts", P),
                ("norm_err", P), ("carter", P), ("g", P), ("dir", P), ("ray_xyz", P)]


def build(cli=False):
    """Compile lib/liboracle.so (and _ref/libbhref.so when the reference is present);
    with cli=True also the reference's program, as shipped and with the GPU
    render() linked in (needs blackhole_amd/lib/libbhx.so built first)."""
    subprocess.run(["make", "-s", "-j8", "-C", HERE, "all"] + (["cli"] if cli else []), check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle library missing: {LIB} (run make -C oracle)")
        _lib = C.CDLL(LIB)
        _lib.oracle_auto_exposure.restype = C.c_double
        _lib.oracle_isco.restype = C.c_double
        _lib.oracle_page_thorne_W.restype = C.c_double
        _lib.oracle_disc_temperature.restype = C.c_double
        _lib.oracle_isco.argtypes = [D, C.c_int]
        _lib.oracle_page_thorne_W.argtypes = [D, D, D]
        _lib.oracle_auto_exposure.argtypes = [P, C.c_int64, D]
        _lib.oracle_develop.argtypes = [P, C.c_int64, D, C.c_int, D, P]
        _lib.oracle_add_glare.argtypes = [P, C.c_int32, C.c_int32, D, D]
        _lib.oracle_blackbody_xyz.argtypes = [D, P]
        _lib.oracle_blackbody_xyz_exact.argtypes = [D, P]
        _lib.oracle_powerlaw_xyz.argtypes = [D, P]
        _lib.oracle_integrate.argtypes = [D, P, D, D, D, C.c_int, D, D, P]
        _lib.oracle_geodesic_rhs.argtypes = [D, P, P]
        _lib.oracle_make_photon.argtypes = [D, D, D, D, P, P]
        _lib.oracle_invariants.argtypes = [D, P, P, P, P, P, P]
        _lib.oracle_star_radiance.argtypes = [P, P, P, D, P]
        _lib.oracle_generate_catalogue.argtypes = [C.c_int32, C.c_uint64, P, P, P, P, P]
    return _lib


def ref_available():
    return os.path.exists(REF)


def ref():
    global _ref
    if _ref is None:
        if not os.path.exists(REF):
            raise RuntimeError(f"reference shim missing: {REF}")
        _ref = C.CDLL(REF)
        _ref.bhref_auto_exposure.restype = C.c_double
        _ref.bhref_auto_exposure.argtypes = [P, C.c_int64, D]
        _ref.bhref_develop.argtypes = [P, C.c_int64, D, C.c_int, D, P]
        _ref.bhref_add_glare.argtypes = [P, C.c_int32, C.c_int32, D, D]
        _ref.bhref_blackbody_xyz.argtypes = [D, P]
        _ref.bhref_star_radiance.argtypes = [C.c_int32, C.c_uint64, P, D, P]
        _ref.bhref_polytrope_star.argtypes = [D, D, D, P, P]
        _ref.bhref_polytrope_solve.argtypes = [D, D, D, P, P]
    return _ref


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


class Cat:
    """Holds numpy arrays alive for a Catalogue struct."""

    def __init__(self, x, y, z, T, irr):
        self.arrays = [np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z, T, irr)]
        self.struct = Catalogue(len(self.arrays[0]), *[_ptr(v) for v in self.arrays])

    @property
    def ref(self):
        return C.byref(self.struct)


def generate_catalogue(count, seed):
    """The reference's StarField(count, seed) catalogue (scene.h:62-107)."""
    arrs = [np.empty(count) for _ in range(5)]
    lib().oracle_generate_catalogue(count, seed, *[_ptr(a) for a in arrs])
    return arrs


def uniform_sky(count, seed):
    """The synthetic uniform sky (the oracle's restatement of bhx_uniform_sky)."""
    arrs = [np.empty(count) for _ in range(5)]
    L = lib()
    L.oracle_uniform_sky.argtypes = [C.c_int32, C.c_uint64, P, P, P, P, P]
    if L.oracle_uniform_sky(count, seed, *[_ptr(a) for a in arrs]) != 0:
        raise ValueError("bad star count")
    return arrs


def render(scene, cat=None, threads=0):
    """Whole-frame render (CPU).  Returns (xyz[H, W, 3], Stats)."""
    xyz = np.zeros((scene.height, scene.width, 3))
    st = Stats()
    rc = lib().oracle_render(C.byref(scene), cat.ref if cat else None, _ptr(xyz), C.byref(st),
                             threads)
    if rc:
        raise RuntimeError(f"oracle_render failed: {rc}")
    return xyz, st


def trace_list(scene, ray_ids, cat=None, threads=0):
    """Trace + shade explicit rays.  Returns a dict of per-ray numpy arrays."""
    ids = np.ascontiguousarray(ray_ids, dtype=np.int64)
    n = len(ids)
    out = {
        "cls": np.zeros(n, np.uint8), "y": np.zeros((8, n)), "jet": np.zeros(n),
        "steps": np.zeros(n, np.int32), "attempts": np.zeros(n, np.int32),
        "norm_err": np.zeros(n), "carter": np.zeros(n), "g": np.zeros(n),
        "dir": np.zeros((3, n)), "ray_xyz": np.zeros((3, n)),
    }
    s = OracleRays(*[_ptr(out[k]) for k in ("cls", "y", "jet", "steps", "attempts", "norm_err",
                                            "carter", "g", "dir", "ray_xyz")])
    rc = lib().oracle_trace_list(C.byref(scene), cat.ref if cat else None, _ptr(ids),
                                 C.c_int64(n), C.byref(s), threads)
    if rc:
        raise RuntimeError(f"oracle_trace_list failed: {rc}")
    return out


def _rays_out(n):
    return {
        "cls": np.zeros(n, np.uint8), "y": np.zeros((8, n)), "jet": np.zeros(n),
        "steps": np.zeros(n, np.int32), "attempts": np.zeros(n, np.int32),
        "norm_err": np.zeros(n), "carter": np.zeros(n), "g": np.zeros(n),
        "dir": np.zeros((3, n)), "ray_xyz": np.zeros((3, n)),
    }


def _rays_struct(out):
    return OracleRays(*[_ptr(out[k]) for k in ("cls", "y", "jet", "steps", "attempts", "norm_err",
                                                "carter", "g", "dir", "ray_xyz")])


def trace_uv(scene, u, v, cat=None, threads=0):
    """Trace + shade rays from explicit image-plane coordinates."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = _rays_out(len(u))
    rc = lib().oracle_trace_uv(C.byref(scene), cat.ref if cat else None, _ptr(u), _ptr(v),
                               C.c_int64(len(u)), C.byref(_rays_struct(out)), threads)
    if rc:
        raise RuntimeError(f"oracle_trace_uv failed: {rc}")
    return out


def ray_uv(scene, ray_ids):
    ids = np.ascontiguousarray(ray_ids, dtype=np.int64)
    u = np.zeros(len(ids))
    v = np.zeros(len(ids))
    lib().oracle_ray_uv(C.byref(scene), _ptr(ids), C.c_int64(len(ids)), _ptr(u), _ptr(v))
    return u, v


def ref_render(scene):
    xyz = np.zeros((scene.height, scene.width, 3))
    st = Stats()
    ref().bhref_render(C.byref(scene), _ptr(xyz), C.byref(st))
    return xyz, st


def ref_trace_list(scene, ray_ids):
    ids = np.ascontiguousarray(ray_ids, dtype=np.int64)
    n = len(ids)
    out = {"cls": np.zeros(n, np.uint8), "y": np.zeros((8, n)), "jet": np.zeros(n),
           "steps": np.zeros(n, np.int32), "g": np.zeros(n), "ray_xyz": np.zeros((3, n))}
    ref().bhref_trace_list(C.byref(scene), _ptr(ids), C.c_int64(n),
                           *[_ptr(out[k]) for k in ("cls", "y", "jet", "steps", "g", "ray_xyz")])
    return out


def jet_band_factor(alpha, nu_0=299792458.0 / 550e-9, nu_lo=1.0e9, nu_hi=1.0e17):
    """kappa = nu_0^alpha * integral of nu^-alpha over [nu_lo, nu_hi]: the
    intensity of a power-law spectrum I_nu = J (nu / nu_0)^-alpha over the
    band, per unit specific intensity J at nu_0 (the band of the conical
    jet's normalisation)."""
    p = 1.0 - alpha
    band = (nu_hi ** p - nu_lo ** p) / p if abs(p) > 1e-9 else math.log(nu_hi / nu_lo)
    return nu_0 ** alpha * band


def disc_intensity(scene, threads=0):
    """Every ray of the scene traced by the oracle, and its intensity: the
    bolometric sigma (g T)^4 / pi of a disc hit (0 elsewhere) plus, with the
    conical jet (jet_model 1), kappa times the jet's specific intensity at
    nu_0 integrated along the ray."""
    ids = np.arange(scene.n_rays)
    c = trace_list(scene, ids, threads=threads)
    cls = np.ascontiguousarray(c["cls"], dtype=np.uint8)
    r = np.ascontiguousarray(c["y"][1])
    g = np.ascontiguousarray(c["g"])
    out = np.zeros(scene.n_rays)
    lib().oracle_disc_intensity(C.byref(scene), C.c_int64(scene.n_rays), _ptr(cls), _ptr(r), _ptr(g), _ptr(out))
    if scene.jet_enabled and scene.jet_model == 1:
        out = out + jet_band_factor(scene.jet_alpha) * c["jet"]
    return out, c


def add_glare(xyz, strength, radius_px=3.0):
    """img::add_glare on an (H, W, 3) image: a new array."""
    x = np.array(xyz, dtype=np.float64, order="C", copy=True)
    H, W = x.shape[:2]
    lib().oracle_add_glare(_ptr(x), W, H, float(strength), float(radius_px))
    return x


def auto_exposure(xyz, key=1.0):
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    return lib().oracle_auto_exposure(_ptr(x), len(x), key)


def develop(xyz, exposure, tonemap=4, log_decades=16.0):
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    rgb = np.zeros((len(x), 3), np.uint8)
    lib().oracle_develop(_ptr(x), len(x), exposure, tonemap, log_decades, _ptr(rgb))
    return rgb


def blackbody_xyz(T):
    o = np.zeros(3)
    lib().oracle_blackbody_xyz(T, _ptr(o))
    return o


def blackbody_xyz_exact(T):
    o = np.zeros(3)
    lib().oracle_blackbody_xyz_exact(T, _ptr(o))
    return o


def geodesic_rhs(a, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    d = np.zeros(8)
    lib().oracle_geodesic_rhs(a, _ptr(y), _ptr(d))
    return d


def make_photon(a, r, th, phi, direction):
    d = np.ascontiguousarray(direction, dtype=np.float64)
    y = np.zeros(8)
    lib().oracle_make_photon(a, r, th, phi, _ptr(d), _ptr(y))
    return y


def invariants(a, y):
    y = np.ascontiguousarray(y, dtype=np.float64)
    vals = [C.c_double() for _ in range(5)]
    lib().oracle_invariants(a, _ptr(y), *[C.byref(v) for v in vals])
    return dict(zip(("E", "Lz", "Q", "norm", "norm_rel"), [v.value for v in vals]))


def integrate(a, y, h0, rtol, atol, max_steps, r_min, r_max):
    y = np.array(y, dtype=np.float64)
    n = C.c_int64(0)
    rc = lib().oracle_integrate(a, _ptr(y), h0, rtol, atol, max_steps, r_min, r_max, C.byref(n))
    return y, rc, n.value


# ---- PNG / APNG writer (src/png.cpp; reference image.cpp:193-778) -----------
def _encode(fn, *args):
    L = lib()
    n = C.c_size_t(0)
    rc = fn(*args, None, C.c_size_t(0), C.byref(n))
    if rc < 0:
        raise ValueError("invalid image")
    out = np.empty(n.value, np.uint8)
    rc = fn(*args, out.ctypes.data_as(P), C.c_size_t(n.value), C.byref(n))
    if rc != 0:
        raise RuntimeError("encoder failed")
    return out.tobytes()


def png_encode(rgb):
    """img::write_png's bytes for an (h, w, 3) uint8 image."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w, _ = rgb.shape
    L = lib()
    L.oracle_png_encode.argtypes = [P, C.c_int32, C.c_int32, P, C.c_size_t, C.POINTER(C.c_size_t)]
    return _encode(L.oracle_png_encode, rgb.ctypes.data_as(P), w, h)


def apng_encode(frames, fps, loops=0):
    """img::write_apng's bytes for an (n, h, w, 3) uint8 stack of frames."""
    frames = np.ascontiguousarray(frames, dtype=np.uint8)
    n, h, w, _ = frames.shape
    L = lib()
    L.oracle_apng_encode.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, C.c_size_t,
                                     C.POINTER(C.c_size_t)]
    return _encode(L.oracle_apng_encode, frames.ctypes.data_as(P), n, w, h, fps, loops)


def huffman_lengths(freq, max_bits):
    """The writer's Huffman code-length builder (image.cpp:270-334)."""
    freq = np.ascontiguousarray(freq, dtype=np.uint32)
    out = np.zeros(len(freq), np.uint8)
    L = lib()
    L.oracle_huffman_lengths.argtypes = [P, C.c_int32, C.c_int32, P]
    L.oracle_huffman_lengths(freq.ctypes.data_as(P), len(freq), max_bits, out.ctypes.data_as(P))
    return out


def _ref_file(call):
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "out.png")
        if call(path.encode()) != 0:
            raise RuntimeError("reference writer failed")
        with open(path, "rb") as f:
            return f.read()


def ref_write_png(rgb):
    """The reference's own img::write_png (oracle/_ref), as bytes."""
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    h, w, _ = rgb.shape
    R = ref()
    R.bhref_write_png.argtypes = [C.c_char_p, P, C.c_int32, C.c_int32]
    return _ref_file(lambda p: R.bhref_write_png(p, rgb.ctypes.data_as(P), w, h))


def ref_write_apng(frames, fps, loops=0):
    """The reference's own img::write_apng (oracle/_ref), as bytes."""
    frames = np.ascontiguousarray(frames, dtype=np.uint8)
    n, h, w, _ = frames.shape
    R = ref()
    R.bhref_write_apng.argtypes = [C.c_char_p, P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    return _ref_file(lambda p: R.bhref_write_apng(p, frames.ctypes.data_as(P), n, w, h, fps, loops))
