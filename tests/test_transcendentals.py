"""The tracer's own sin/cos, exp and log (blackhole_amd/csrc/kerr_dev.h,
used instead of the device library's) against glibc, on the CPU: a host build
of the same code over 2M arguments each (theta-like angles, wide angles up to
1e5, tiny angles; exp over [-700, 700]; log over the normal range) must be
within 1 ulp of glibc, which the reference uses.  Next to the zeros of sin
and cos (x a multiple of pi/4 to within 1e-12, |x| <= 200) the two-term
argument reduction leaves an absolute error of at most 1 ulp + |k| 4e-33
(k = x / (pi/2) rounded): relative error grows there as the value vanishes
(451 ulps at x = 29 pi/2, cos x = -6.2e-19), but no sign the tracer tests can
change (the equatorial crossing test would need |cos theta| < 1e-28; the
double nearest pi/2 has cos = 6.1e-17)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_within_one_ulp_of_glibc(tmp_path):
    if not os.path.exists(HIPCC) or shutil.which("g++") is None:
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "check")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", os.path.join(HERE, "transcendentals", "check.cpp"), "-o", exe],
                   check=True, capture_output=True)
    out = subprocess.run([exe, "2000000"], check=True, capture_output=True, text=True, timeout=300).stdout
    res = {ln.split()[0]: (int(ln.split()[1]), int(ln.split()[2])) for ln in out.strip().splitlines()}
    print(res)
    for name in ("sin", "cos", "exp", "log"):
        n, worst = res[name]
        assert n >= 2000000 and worst <= 1, (name, worst)
    assert res["boundary"] == (1539, 0)
