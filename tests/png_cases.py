"""Deterministic test images for the PNG / APNG writer (shared by the tests
and tests/golden/make_golden_png.py, which records the reference writer's
bytes for each)."""
import numpy as np


def _disc(w, h, seed):
    """A dark sky with a smooth tilted ring and a few hundred point stars."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    u = (x - 0.5 * w) / (0.45 * w)
    v = (y - 0.5 * h) / (0.18 * h)
    r = np.sqrt(u * u + v * v)
    ring = np.clip(1.0 - np.abs(r - 0.7) / 0.3, 0.0, 1.0) ** 1.5
    img = np.zeros((h, w, 3))
    img[..., 0] = 255.0 * ring
    img[..., 1] = 190.0 * ring ** 1.4
    img[..., 2] = 120.0 * ring ** 2.2
    img[r < 0.35] = 0.0
    n = max(1, (w * h) // 400)
    sx, sy = rng.integers(0, w, n), rng.integers(0, h, n)
    img[sy, sx] = rng.integers(40, 256, (n, 3))
    return np.round(img).astype(np.uint8)


def images():
    """name -> (h, w, 3) uint8."""
    rng = np.random.default_rng(7)
    y, x = np.mgrid[0:23, 0:37]
    out = {
        "gradient_37x23": np.stack([(x * 7) % 256, (y * 11) % 256, (x * y) % 256], -1).astype(np.uint8),
        "one_pixel": np.array([[[200, 10, 30]]], np.uint8),
        "noise_64x48": rng.integers(0, 256, (48, 64, 3), dtype=np.uint8),
        "black_200x100": np.zeros((100, 200, 3), np.uint8),
        "wide_1000x3": np.repeat(((np.arange(1000) * 3) % 256).astype(np.uint8)[None, :, None], 3, 0).repeat(3, 2),
        "tall_3x500": np.stack([np.arange(500) % 256] * 9, -1).reshape(500, 3, 3).astype(np.uint8),
        "disc_160x120": _disc(160, 120, 1),
        "disc_512x384": _disc(512, 384, 2),
    }
    return out


def animations():
    """name -> (frames (n, h, w, 3) uint8, fps, loops)."""
    base = _disc(64, 48, 3)
    frames = []
    for k in range(5):
        f = base.copy()
        if k < 3:
            f[20:23, 10 + 6 * k:13 + 6 * k] = (255, 255, 200)   # a spot that moves
        else:
            f[20:23, 22:25] = (255, 255, 200)
        frames.append(f)
    frames[3] = frames[2].copy()                                  # identical frame: a 1x1 sub-frame
    frames[4] = (frames[2].astype(np.int32) + 1).clip(0, 255).astype(np.uint8)   # every pixel changes
    spot = np.stack(frames)
    return {"spot_64x48": (spot, 24, 0), "spot_fps300_loop3": (spot[:3], 300, 3)}
