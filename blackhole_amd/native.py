"""ctypes binding of libbhx.so (the HIP tracer's C ABI, include/bhx.h).

There is no CPU fallback: if the library is missing or cannot be loaded this
module raises, and every entry point that needs the GPU fails loudly.
"""
import ctypes as C
import os

from .abi import Bytes, Launch, Rays, Scene, Shaded, Stats

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libbhx.so")

_lib = None

P = C.c_void_p
SYMBOLS = (
    "bhx_create", "bhx_destroy", "bhx_last_error", "bhx_abi_version", "bhx_set_catalogue",
    "bhx_render", "bhx_trace_dev", "bhx_shade_dev", "bhx_auto_exposure_dev", "bhx_develop_dev", "bhx_develop_auto_dev", "bhx_add_glare_dev",
    "bhx_blackbody_xyz_dev", "bhx_set_launch", "bhx_get_launch", "bhx_setup_jet",
    "bhx_polytrope_star", "bhx_setup_neutron_star", "bhx_setup_conical_jet", "bhx_batch_create", "bhx_batch_destroy",
    "bhx_batch_run", "bhx_batch_scene", "bhx_batch_copy_scene", "bhx_batch_stats",
    "bhx_batch_set_launch", "bhx_batch_get_launch", "bhx_batch_autotune", "bhx_adjoint_create",
    "bhx_adjoint_destroy", "bhx_adjoint_set_scene", "bhx_adjoint_forward", "bhx_adjoint_backward",
    "bhx_adjoint_rays", "bhx_adjoint_copy_rays", "bhx_adjoint_info", "bhx_png_encode", "bhx_png_encode_dev",
    "bhx_apng_encode", "bhx_apng_encode_dev", "bhx_bytes_free", "bhx_trace_noalloc_dev", "bhx_uniform_sky",
    "bhx_adjoint_autotune", "bhx_adjoint_get_interval", "bhx_adjoint_create_batch", "bhx_adjoint_backward_batch",
    "bhx_adjoint_set_launch", "bhx_adjoint_set_forward_waves", "bhx_adjoint_get_forward_waves",
)


class BhxError(RuntimeError):
    pass


def lib():
    """Load libbhx.so once (no GPU call is made by loading)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BhxError(f"HIP library not built: {LIB_PATH} (run python -c "
                       f"'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(LIB_PATH)
    L.bhx_create.argtypes = [C.c_int, C.POINTER(P)]
    L.bhx_destroy.argtypes = [P]
    L.bhx_last_error.argtypes = [P]
    L.bhx_last_error.restype = C.c_char_p
    L.bhx_abi_version.restype = C.c_int
    L.bhx_set_catalogue.argtypes = [P, C.c_int32, P, P, P, P, P]
    L.bhx_render.argtypes = [P, C.POINTER(Scene), P, C.POINTER(Stats)]
    L.bhx_trace_dev.argtypes = [P, C.POINTER(Scene), C.POINTER(Rays), C.POINTER(Stats), P]
    L.bhx_shade_dev.argtypes = [P, C.POINTER(Scene), C.POINTER(Rays), C.POINTER(Shaded),
                                C.POINTER(Stats), P]
    L.bhx_auto_exposure_dev.argtypes = [P, P, C.c_int64, C.c_double, C.POINTER(C.c_double), P]
    L.bhx_develop_dev.argtypes = [P, P, C.c_int64, C.c_double, C.c_int32, C.c_double, P, P]
    L.bhx_develop_auto_dev.argtypes = [P, P, C.c_int64, C.c_double, C.c_int32, C.c_double, P, P, P]
    L.bhx_add_glare_dev.argtypes = [P, P, C.c_int32, C.c_int32, C.c_double, C.c_double, P]
    L.bhx_blackbody_xyz_dev.argtypes = [P, P, C.c_int64, P, P]
    L.bhx_set_launch.argtypes = [P, C.POINTER(Launch)]
    L.bhx_get_launch.argtypes = [P, C.POINTER(Launch)]
    L.bhx_setup_jet.argtypes = [C.POINTER(Scene), C.c_double, C.c_double]
    L.bhx_polytrope_star.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
    L.bhx_setup_neutron_star.argtypes = [C.POINTER(Scene), C.c_double, C.c_double]
    L.bhx_uniform_sky.argtypes = [C.c_int32, C.c_uint64, P, P, P, P, P]
    L.bhx_setup_conical_jet.argtypes = [C.POINTER(Scene), C.c_double, C.c_double]
    L.bhx_batch_create.argtypes = [P, C.POINTER(Scene), C.c_int32, C.POINTER(P)]
    L.bhx_batch_destroy.argtypes = [P]
    L.bhx_batch_run.argtypes = [P, C.c_int32, P]
    L.bhx_batch_scene.argtypes = [P, C.c_int32, C.POINTER(Rays), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), C.POINTER(P)]
    L.bhx_batch_copy_scene.argtypes = [P, C.c_int32, P, P, P, P, P, P, P]
    L.bhx_batch_stats.argtypes = [P, C.POINTER(Stats), C.POINTER(C.c_double), P]
    L.bhx_batch_set_launch.argtypes = [P, C.POINTER(Launch)]
    L.bhx_batch_get_launch.argtypes = [P, C.POINTER(Launch)]
    L.bhx_batch_autotune.argtypes = [P, C.c_int32, C.POINTER(Launch), C.POINTER(C.c_double), P]
    L.bhx_adjoint_create.argtypes = [P, C.POINTER(Scene), C.c_int32, C.POINTER(P)]
    L.bhx_adjoint_destroy.argtypes = [P]
    L.bhx_adjoint_create_batch.argtypes = [P, C.POINTER(Scene), C.c_int32, C.c_int32, C.POINTER(P)]
    L.bhx_adjoint_backward_batch.argtypes = [P, P, P, P]
    L.bhx_adjoint_set_launch.argtypes = [P, C.c_int32, C.c_int32]
    L.bhx_adjoint_set_forward_waves.argtypes = [P, C.c_int32]
    L.bhx_adjoint_get_forward_waves.argtypes = [P, C.POINTER(C.c_int32)]
    L.bhx_adjoint_set_scene.argtypes = [P, C.POINTER(Scene)]
    L.bhx_adjoint_forward.argtypes = [P, P, P]
    L.bhx_adjoint_backward.argtypes = [P, P, P, P]
    L.bhx_adjoint_rays.argtypes = [P, C.POINTER(Rays)]
    L.bhx_adjoint_copy_rays.argtypes = [P, P, P, P, P]
    L.bhx_adjoint_autotune.argtypes = [P, P, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_double), P]
    L.bhx_adjoint_get_interval.argtypes = [P, C.POINTER(C.c_int32)]
    L.bhx_adjoint_info.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.bhx_png_encode.argtypes = [P, P, C.c_int32, C.c_int32, C.POINTER(Bytes)]
    L.bhx_png_encode_dev.argtypes = [P, P, C.c_int32, C.c_int32, P, C.POINTER(Bytes)]
    L.bhx_apng_encode.argtypes = [P, C.POINTER(P), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.POINTER(Bytes)]
    L.bhx_apng_encode_dev.argtypes = [P, C.POINTER(P), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P,
                                      C.POINTER(Bytes)]
    L.bhx_trace_noalloc_dev.argtypes = [C.c_int32, C.POINTER(Scene), C.POINTER(Rays), C.POINTER(Stats), P]
    L.bhx_bytes_free.argtypes = [C.POINTER(Bytes)]
    L.bhx_bytes_free.restype = None
    _lib = L
    return L


def check(ctx, rc, what):
    if rc != 0:
        msg = lib().bhx_last_error(ctx)
        raise BhxError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
