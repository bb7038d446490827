"""blackhole_amd - MI355X-native (gfx950, HIP, FP64) Kerr geodesic tracer.

A drop-in for the per-pixel tracer of MattPar-Cell/Blackhole: the reference's
frame API bh::render(RenderConfig, RenderStats) (src/render.h:72) is kept, and
everything below it - Dormand-Prince integration of the Kerr Hamiltonian flow,
event location, disc / neutron-star / sky shading, colour - runs in the HIP
library libbhx.so through the C ABI of include/bhx.h.

    from blackhole_amd import Tracer, scene
    with Tracer() as tr:
        xyz, stats = tr.render(scene.plain_frame())
"""
from . import abi, scene
from .abi import CLASS_NAMES, Scene, Stats
from .native import BhxError

__all__ = ["abi", "scene", "Scene", "Stats", "CLASS_NAMES", "BhxError", "Tracer", "Batch"]


def __getattr__(name):
    # Tracer imports torch lazily so that `import blackhole_amd` stays cheap.
    if name in ("Tracer", "Batch"):
        from . import tracer
        return getattr(tracer, name)
    raise AttributeError(name)
