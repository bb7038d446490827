"""The reference's gallery, rendered by the drop-in.

The reference ships 32 stills (gallery/*.png) and the command lines that
rendered them (Makefile:33-110): black holes at up to 3840x2160 with 2-3
samples per pixel, the Sun, conical jets and neutron stars from a TOV solve.
The reference program built with src/render.cpp replaced by
blackhole_amd/integration/render_hip.cpp and its PNG writer routed to
blackhole_amd/integration/png_hip.cpp (oracle/Makefile target `cli`) runs each
command line here, and its file must be the reference's file byte for byte:
tests/golden/gallery.json holds each image's length and SHA-256.  Needs an
MI355X; skipped where the drop-in program was not built."""
import hashlib
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "blackhole_hip")
with open(os.path.join(HERE, "golden", "gallery.json")) as f:
    CASES = json.load(f)["cases"]


@pytest.mark.skipif(not os.path.exists(HIP_BIN),
                    reason="drop-in program not built (make -C oracle cli, needs the reference sources)")
@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gallery_image_reproduced(tmp_path, case):
    out = tmp_path / case["name"]
    cmd = [HIP_BIN, *case["args"], "--out", str(out)]
    print(f"[gallery {case['name']}] {' '.join(case['args'])}", flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, (cmd, p.stdout[-2000:], p.stderr[-2000:])
    data = out.read_bytes()
    assert len(data) == case["bytes"]
    assert hashlib.sha256(data).hexdigest() == case["sha256"]
