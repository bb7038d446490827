"""The tracer stage without device allocation (bhx_trace_noalloc_dev) against
bhx_trace_dev, through the C ABI.

Both write the same per-ray outputs - class, hit state, jet integral,
accepted steps, attempts, invariant maxima - bit for bit, and the same stats
counters.  Checked on calls in a row whose frames shrink (the scratch array
keeps the last call's contents), on a conical-jet frame (event parameters and
the spin-axis queue band in the scratch array), and on a frame at the size limit
(the refusal of larger frames is host logic, tested on the CPU in
test_abi).  Output buffers start filled with a
sentinel so that a ray the call did not write shows up.
"""
import ctypes as C

import pytest

from blackhole_amd import native
from blackhole_amd import scene as S
from blackhole_amd.abi import Stats

pytestmark = pytest.mark.gpu

KEYS = ("cls", "y", "jet", "steps", "attempts", "norm_err", "carter")
COUNTERS = ("rays", "steps", "attempts", "rhs_evals", "max_norm_error", "max_carter_drift")


def _sentinel_rays(tracer, sc):
    import torch
    r = tracer.alloc_rays(sc)
    for t in r.values():
        if t.dtype == torch.uint8:
            t.fill_(0xAB)
        elif t.dtype == torch.int32:
            t.fill_(-7)
        else:
            t.fill_(float("nan"))
    return r


def _with_context(tracer, sc):
    rays = _sentinel_rays(tracer, sc)
    _, st = tracer.trace(sc, rays=rays)
    return rays, st


def _without_allocation(tracer, sc):
    rays = _sentinel_rays(tracer, sc)
    st = Stats()
    rc = native.lib().bhx_trace_noalloc_dev(tracer.device, C.byref(sc), C.byref(tracer._rays_struct(rays)),
                                            C.byref(st), tracer._stream(None))
    assert rc == 0, native.lib().bhx_last_error(None)
    return rays, st


def _same_bits(a, b):
    import torch
    if a.dtype == torch.float64:
        return torch.equal(a.view(torch.int64), b.view(torch.int64))
    return torch.equal(a, b)


def _check(tracer, sc, label):
    print(f"[no-alloc] {label}: {sc.width}x{sc.height} spp {sc.sqrt_spp ** 2}", flush=True)
    ra, sa = _with_context(tracer, sc)
    rb, sb = _without_allocation(tracer, sc)
    for k in KEYS:
        assert _same_bits(ra[k], rb[k]), (label, k)
    for k in COUNTERS:
        assert getattr(sa, k) == getattr(sb, k), (label, k, getattr(sa, k), getattr(sb, k))
    assert sb.rays == sc.n_rays and sb.steps > 0


def test_two_calls_in_a_row_second_smaller(tracer):
    _check(tracer, S.plain_frame(96, 80), "plain")
    sc = S.starry_frame(40, 24)
    sc.sqrt_spp = 2
    _check(tracer, sc, "starry, 2x2 samples")


def test_conical_jet_frame(tracer):
    sc = S.quasar_frame(10.0, 48, 40, stars=False, half_extent=120.0)
    _check(tracer, sc, "conical jet")


def test_frame_at_the_limit(tracer):
    _check(tracer, S.starry_frame(1024, 1024), "1024^2 starry")
    _check(tracer, S.plain_frame(33, 17), "33x17 after the limit frame")

