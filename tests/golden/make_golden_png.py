"""Record the reference writer's PNG / APNG bytes for the test images.

Runs img::write_png / img::write_apng of the reference compiled from its own
sources (oracle/_ref/libbhref.so, see oracle/Makefile) on the deterministic
images of tests/png_cases.py and writes tests/golden/png_reference.json: per
case the stream's length and SHA-256 (the data is the fixture; the images are
regenerated from their seeds).
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import oracle  # noqa: E402
from png_cases import animations, images  # noqa: E402


def digest(b):
    return {"bytes": len(b), "sha256": hashlib.sha256(b).hexdigest()}


def main():
    if not oracle.ref_available():
        raise SystemExit("reference shim not built: make -C oracle (needs /root/reference)")
    out = {"note": "written by tests/golden/make_golden_png.py from oracle/_ref (img::write_png / write_apng)",
           "png": {}, "apng": {}}
    for name, rgb in images().items():
        out["png"][name] = digest(oracle.ref_write_png(rgb))
    for name, (frames, fps, loops) in animations().items():
        out["apng"][name] = dict(digest(oracle.ref_write_apng(frames, fps, loops)), fps=fps, loops=loops)
    with open(os.path.join(HERE, "png_reference.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
