"""The adjoint's reverse mode against central finite differences of the same
forward code, on the CPU: tests/adjoint/selfcheck.cpp built with hipcc as a
host program (the device functions are host-callable).  Checks the
right-hand side's vector-Jacobian product, one frozen-step Dormand-Prince
step, the terminal step's implicit-function-theorem crossing, and whole
disc-hit rays against finite differences of the full adaptive forward."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def selfcheck(tmp_path_factory):
    if not os.path.exists(HIPCC) or shutil.which("g++") is None:
        pytest.skip("hipcc / g++ not available")
    d = tmp_path_factory.mktemp("adjcheck")
    hs = str(d / "host_setup.o")
    subprocess.run(["g++", "-O2", "-std=c++17", "-c", os.path.join(ROOT, "blackhole_amd", "csrc", "host_setup.cpp"),
                    "-o", hs], check=True)
    obj = str(d / "selfcheck.o")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-c", os.path.join(HERE, "adjoint", "selfcheck.cpp"), "-o", obj],
                   check=True, capture_output=True)
    exe = str(d / "selfcheck")
    subprocess.run([HIPCC, obj, hs, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600).stdout
    print(out)
    rows = {}
    for ln in out.splitlines():
        parts = ln.split()
        if len(parts) >= 3 and not ln.startswith(" "):
            rows[parts[0]] = {"n": int(parts[1]), "err": float(parts[2]),
                              "mismatch": int(parts[4]) if len(parts) > 4 else 0}
    return rows


def test_rhs_vjp(selfcheck):
    assert selfcheck["rhs"]["n"] >= 100 and selfcheck["rhs"]["err"] < 1e-5


def test_step_vjp(selfcheck):
    assert selfcheck["step"]["n"] >= 50 and selfcheck["step"]["err"] < 1e-4


def test_terminal_step_ift(selfcheck):
    assert selfcheck["term"]["n"] >= 50 and selfcheck["term"]["err"] < 1e-5


def test_whole_rays_against_adaptive_fd(selfcheck):
    for k in ("ray_rtol1e-09_K16", "ray_rtol1e-12_K7"):
        r = selfcheck[k]
        assert r["n"] >= 20 and r["mismatch"] == 0, (k, r)
    assert selfcheck["ray_rtol1e-09_K16"]["err"] < 1e-4
    assert selfcheck["ray_rtol1e-12_K7"]["err"] < 1e-6
