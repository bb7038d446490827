"""Cost-ordered work queue (pilot + sort) vs the centre-first queue (needs an MI355X).

The queue order decides only when each 8 x 8 tile starts; every ray is
integrated on its own, so both orders must give the same per-ray results to
the bit.  Output buffers are pre-filled with a sentinel so that a tile left
out of (or traced twice in) the sorted queue shows up as a mismatch.
"""
import pytest

from blackhole_amd import scene as S
from blackhole_amd.abi import ORDER_CENTRE, ORDER_COST

pytestmark = pytest.mark.gpu

KEYS = ("cls", "y", "jet", "steps", "attempts", "norm_err", "carter")


def _trace(tracer, sc, order):
    tracer.set_launch(queue_order=order)
    rays = tracer.alloc_rays(sc)
    for k, v in rays.items():
        v.fill_(255 if k == "cls" else -7)
    return tracer.trace(sc, rays=rays)


@pytest.mark.parametrize("shape", [(1024, 1024), (1000, 700)])
def test_cost_order_traces_the_same_rays(tracer, shape):
    torch = tracer.torch
    w, h = shape
    sc = S.starry_frame(w, h, stars=False)
    print(f"[queue order] {w}x{h} a=0.998 i=60 deg, cost vs centre-first", flush=True)
    try:
        rc, stc = _trace(tracer, sc, ORDER_COST)
        rr, str_ = _trace(tracer, sc, ORDER_CENTRE)
    finally:
        tracer.set_launch(queue_order=ORDER_COST)
    print(f"[queue order] cost: trace {stc.trace_ms:.3f} ms + order {stc.order_ms:.3f} ms; "
          f"centre-first: trace {str_.trace_ms:.3f} ms", flush=True)
    assert stc.order_ms > 0.0 and str_.order_ms == 0.0   # the pilot ran only for the cost order
    assert int((rc["cls"] == 255).sum()) == 0             # every ray written
    for k in KEYS:   # bit patterns (a NaN equals itself here)
        a, b = rc[k], rr[k]
        if a.dtype == torch.float64:
            a, b = a.view(torch.int64), b.view(torch.int64)
        assert torch.equal(a, b), k
    assert (stc.steps, stc.attempts, stc.rays) == (str_.steps, str_.attempts, str_.rays)
    assert stc.max_norm_error == str_.max_norm_error and stc.max_carter_drift == str_.max_carter_drift


def test_small_frame_keeps_centre_order(tracer):
    # 256 x 256 = 65,536 rays is fewer than 4 per resident lane: no pilot
    sc = S.plain_frame()
    _, st = tracer.trace(sc)
    assert st.order_ms == 0.0


def test_back_to_back_frames_with_the_order_on_its_own_stream(tracer):
    """Frames queued back to back without a host synchronisation: each
    frame's cost order is built on the context's order stream while the
    previous frame still traces, and must be that frame's own (same rays, to
    the bit, as the frame traced alone).  Three scene shapes in turn, twice
    over, so an order built for the wrong frame or read after it was replaced
    would show."""
    torch = tracer.torch
    tracer.set_launch(queue_order=ORDER_COST)
    scenes = [S.starry_frame(1024, 1024, stars=False), S.starry_frame(1000, 700, stars=False),
              S.starry_frame(1024, 600, stars=False)]
    alone = []
    for sc in scenes:
        rays, st = tracer.trace(sc)
        assert st.order_ms > 0.0   # every shape here is cost-ordered
        alone.append(rays)
    queued = []
    for rep in range(2):
        for sc in scenes:
            rays = tracer.alloc_rays(sc)
            for k, v in rays.items():
                v.fill_(255 if k == "cls" else -7)
            tracer.trace_async(sc, rays)
            queued.append(rays)
    torch.cuda.synchronize()
    for i, rays in enumerate(queued):
        ref = alone[i % len(scenes)]
        for k in KEYS:
            assert torch.equal(rays[k], ref[k]), (i, k)
    print(f"[queue order] {len(queued)} back-to-back frames equal to the frames traced alone", flush=True)
