"""GPU-assisted PNG / APNG writer (blackhole_amd/csrc/png.hip) against the oracle
and the reference writer's bytes (needs an MI355X).

The bar is byte identity: every stream must equal the oracle's (itself equal
to the reference's img::write_png / write_apng, tests/test_png_oracle.py) and,
for the golden cases, the reference's recorded SHA-256.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from blackhole_amd import scene as S
from png_cases import animations, images

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "png_reference.json")))


def sha(b):
    return hashlib.sha256(b).hexdigest()


def test_png_cases_host_and_device_input(tracer):
    torch = tracer.torch
    for name, rgb in images().items():
        print(f"[png] {name} {rgb.shape[1]}x{rgb.shape[0]}", flush=True)
        b = tracer.png_encode(rgb)
        assert sha(b) == GOLDEN["png"][name]["sha256"], name
        bd = tracer.png_encode(torch.from_numpy(rgb).to("cuda:0"))
        assert bd == b, name


def test_apng_cases(tracer):
    torch = tracer.torch
    for name, (frames, fps, loops) in animations().items():
        print(f"[apng] {name} {len(frames)} frames", flush=True)
        b = tracer.apng_encode(list(frames), fps, loops)
        assert sha(b) == GOLDEN["apng"][name]["sha256"], name
        dev = [torch.from_numpy(np.ascontiguousarray(f)).to("cuda:0") for f in frames]
        assert tracer.apng_encode(dev, fps, loops) == b, name


def test_png_random_images_match_oracle(tracer):
    rng = np.random.default_rng(5)
    for k in range(10):
        h, w = int(rng.integers(1, 300)), int(rng.integers(1, 300))
        kind = k % 3
        if kind == 0:
            img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        elif kind == 1:
            img = (rng.integers(0, 3, (h, w, 3)) * 60).astype(np.uint8)
        else:
            y, x = np.mgrid[0:h, 0:w]
            img = np.stack([x * 3 + y, x ^ y, (x * y) >> 2], -1).astype(np.uint8)
        print(f"[png] random {k} {w}x{h} kind {kind}", flush=True)
        assert tracer.png_encode(img) == oracle.png_encode(img), (k, w, h)


def test_png_of_a_rendered_frame(tracer):
    """The display path end to end at 1024^2: render, expose, develop on the
    device, encode from device memory; the same bytes as the oracle's writer
    on the same pixels, and the pixels decode back."""
    import zlib
    sc = S.starry_frame(1024, 1024)
    print("[png] starry 1024x1024 frame", flush=True)
    rays, st = tracer.trace(sc)
    sh, st = tracer.shade(sc, rays, stats=st)
    e = tracer.auto_exposure(sh["xyz"])
    rgb = tracer.develop(sh["xyz"], e).view(sc.height, sc.width, 3)
    b = tracer.png_encode(rgb)
    host = rgb.cpu().numpy()
    assert b == oracle.png_encode(host)
    # IDAT decompresses to (3 W + 1) H filtered bytes
    i = b.index(b"IDAT")
    n = int.from_bytes(b[i - 4:i], "big")
    raw = zlib.decompress(b[i + 4:i + 4 + n])
    assert len(raw) == (3 * sc.width + 1) * sc.height


def test_png_rejects_bad_sizes(tracer):
    from blackhole_amd import native
    with pytest.raises(native.BhxError):
        tracer.png_encode(np.zeros((0, 5, 3), np.uint8))
