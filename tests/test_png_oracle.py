"""The oracle's PNG / APNG writer (oracle/src/png.cpp) against the reference's
img::write_png / write_apng (image.cpp:193-778), and against the format itself.

* golden: the reference writer's bytes (SHA-256, length) for the images of
  tests/png_cases.py, recorded by tests/golden/make_golden_png.py;
* live: where the reference build exists (oracle/_ref), more images compared
  byte for byte;
* decode: every stream is parsed independently (chunk CRCs, zlib, scanline
  unfiltering, APNG frame composition with the fcTL rectangles) and must give
  the input pixels back;
* the Huffman builder: lengths within the limit, a complete prefix code, and
  the optimal cost whenever no flattening was needed.
"""
import hashlib
import heapq
import json
import os
import struct
import zlib

import numpy as np
import pytest

import oracle
from png_cases import animations, images

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "png_reference.json")))


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("name", sorted(GOLDEN["png"]))
def test_png_golden(name):
    b = oracle.png_encode(images()[name])
    assert (len(b), sha(b)) == (GOLDEN["png"][name]["bytes"], GOLDEN["png"][name]["sha256"])


@pytest.mark.parametrize("name", sorted(GOLDEN["apng"]))
def test_apng_golden(name):
    frames, fps, loops = animations()[name]
    b = oracle.apng_encode(frames, fps, loops)
    assert (len(b), sha(b)) == (GOLDEN["apng"][name]["bytes"], GOLDEN["apng"][name]["sha256"])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) not present")
def test_png_live_against_reference():
    rng = np.random.default_rng(11)
    for k in range(12):
        h, w = int(rng.integers(1, 90)), int(rng.integers(1, 130))
        kind = k % 3
        if kind == 0:
            img = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        elif kind == 1:   # few distinct values: long matches, skewed codes
            img = (rng.integers(0, 3, (h, w, 3)) * 60).astype(np.uint8)
        else:
            y, x = np.mgrid[0:h, 0:w]
            img = np.stack([x * 3 + y, x ^ y, (x * y) >> 2], -1).astype(np.uint8)
        assert oracle.png_encode(img) == oracle.ref_write_png(img), (k, h, w)
    frames = (rng.integers(0, 2, (4, 20, 30, 3)) * 255).astype(np.uint8)
    frames[2] = frames[1]
    assert oracle.apng_encode(frames, 7, 1) == oracle.ref_write_apng(frames, 7, 1)


# ---- an independent decoder --------------------------------------------------
def chunks(b):
    assert b[:8] == bytes([137, 80, 78, 71, 13, 10, 26, 10])
    i, out = 8, []
    while i < len(b):
        n = struct.unpack(">I", b[i:i + 4])[0]
        tag, body = b[i + 4:i + 8], b[i + 8:i + 8 + n]
        crc = struct.unpack(">I", b[i + 8 + n:i + 12 + n])[0]
        assert crc == zlib.crc32(tag + body), tag
        out.append((tag.decode(), body))
        i += 12 + n
    assert out[-1][0] == "IEND"
    return out


def unfilter(raw, w, h):
    stride = 3 * w
    img = np.zeros((h, stride), np.int32)
    for y in range(h):
        t, row = raw[y * (stride + 1)], np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        up = img[y - 1] if y else np.zeros(stride, np.int32)
        cur = np.zeros(stride, np.int32)
        for x in range(stride):
            a = cur[x - 3] if x >= 3 else 0
            b = up[x]
            c = up[x - 3] if x >= 3 else 0
            if t == 0:
                p = 0
            elif t == 1:
                p = a
            elif t == 2:
                p = b
            elif t == 3:
                p = (a + b) // 2
            else:
                q = a + b - c
                pa, pb, pc = abs(q - a), abs(q - b), abs(q - c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[x] = (row[x] + p) & 255
        img[y] = cur
    return img.reshape(h, w, 3).astype(np.uint8)


def decode_zlib(z, w, h):
    raw = zlib.decompress(z)
    assert len(raw) == (3 * w + 1) * h
    return unfilter(raw, w, h)


@pytest.mark.parametrize("name", ["gradient_37x23", "one_pixel", "noise_64x48", "black_200x100", "disc_160x120"])
def test_png_decodes_to_the_image(name):
    img = images()[name]
    c = chunks(oracle.png_encode(img))
    assert [t for t, _ in c] == ["IHDR", "IDAT", "IEND"]
    w, h, depth, ctype = struct.unpack(">IIBB", c[0][1][:10])
    assert (w, h, depth, ctype) == (img.shape[1], img.shape[0], 8, 2)
    assert np.array_equal(decode_zlib(c[1][1], w, h), img)


@pytest.mark.parametrize("name", sorted(GOLDEN["apng"]))
def test_apng_decodes_to_the_frames(name):
    frames, fps, loops = animations()[name]
    c = chunks(oracle.apng_encode(frames, fps, loops))
    n, lp = struct.unpack(">II", dict(c)["acTL"])
    assert (n, lp) == (len(frames), loops)
    h, w = frames.shape[1:3]
    canvas = np.zeros((h, w, 3), np.uint8)
    seq, k, rect = 0, 0, None
    for tag, body in c:
        if tag == "fcTL":
            s, fw, fh, fx, fy, dn, dd, dop, bop = struct.unpack(">IIIIIHHBB", body)
            assert s == seq and (dn, dd, dop, bop) == (1, fps, 0, 0)
            seq += 1
            rect = (fx, fy, fw, fh)
        elif tag in ("IDAT", "fdAT"):
            if tag == "fdAT":
                assert struct.unpack(">I", body[:4])[0] == seq
                seq += 1
                body = body[4:]
            fx, fy, fw, fh = rect
            canvas[fy:fy + fh, fx:fx + fw] = decode_zlib(body, fw, fh)
            assert np.array_equal(canvas, frames[k]), k
            k += 1
    assert k == len(frames)


# ---- the Huffman builder -------------------------------------------------------
def optimal_cost(freq):
    f = [int(v) for v in freq if v]
    if len(f) < 2:
        return None
    heapq.heapify(f)
    cost = 0
    while len(f) > 1:
        a, b = heapq.heappop(f), heapq.heappop(f)
        cost += a + b
        heapq.heappush(f, a + b)
    return cost


@pytest.mark.parametrize("seed", range(6))
def test_huffman_lengths(seed):
    rng = np.random.default_rng(seed)
    n, max_bits = (288, 15) if seed % 2 == 0 else (19, 7)
    if seed < 2:        # Fibonacci weights: the unlimited tree is far too deep
        fib = [1, 1]
        while len(fib) < n:
            fib.append(fib[-1] + fib[-2])
        freq = np.array([min(v, 2 ** 31) for v in fib[:n]], np.uint64).astype(np.uint32)
    else:
        freq = rng.integers(0, 1000, n).astype(np.uint32) * (rng.random(n) < 0.6)
    lens = oracle.huffman_lengths(freq, max_bits).astype(np.int64)
    used = freq > 0
    assert lens.max() <= max_bits
    assert np.all(lens[used] > 0) and np.all(lens[~used] == 0)
    kraft = sum(2.0 ** -int(v) for v in lens if v)
    assert kraft == pytest.approx(1.0)            # a complete prefix code
    opt = optimal_cost(freq)
    if seed >= 2 and opt is not None and max(lens) < max_bits:
        assert int((freq.astype(np.int64) * lens).sum()) == opt


def test_huffman_single_symbol():
    for s in (0, 5):
        f = np.zeros(30, np.uint32)
        f[s] = 9
        lens = oracle.huffman_lengths(f, 15)
        other = 1 if s == 0 else 0
        assert lens[s] == 1 and lens[other] == 1 and lens.sum() == 2
