"""The adjoint's reverse mode against central finite differences of the same
forward code, on the CPU: tests/adjoint/selfcheck.cpp built with hipcc as a
host program (the device functions are host-callable).  Checks the
right-hand side's vector-Jacobian product, one frozen-step Dormand-Prince
step, the terminal step's implicit-function-theorem crossing, and whole
rays against finite differences of the full adaptive forward: disc hits, and
in a scene with the conical jet the rays the jet lights (its path integral,
cut at the disc crossing where the ray ends there)."""
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
    objs = []
    for src in ("host_setup.cpp", "host_physics.cpp"):
        o = str(d / (src[:-4] + ".o"))
        subprocess.run(["g++", "-O2", "-std=c++17", "-c", os.path.join(ROOT, "blackhole_amd", "csrc", src), "-o", o],
                       check=True)
        objs.append(o)
    obj = str(d / "selfcheck.o")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", "-c", os.path.join(HERE, "adjoint", "selfcheck.cpp"), "-o", obj],
                   check=True, capture_output=True)
    exe = str(d / "selfcheck")
    subprocess.run([HIPCC, obj, *objs, "-o", exe], check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=600).stdout
    print(out)
    rows = {}
    for ln in out.splitlines():
        parts = ln.split()
        if len(parts) >= 3 and not ln.startswith(" "):
            rows[parts[0]] = {"n": int(parts[1]), "err": float(parts[2]),
                              "mismatch": int(parts[4]) if len(parts) > 4 else 0,
                              "jet": int(parts[6]) if len(parts) > 6 else 0}
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


def test_whole_rays_with_conical_jet(selfcheck):
    # The jet's 8 x 4-point quadrature per step is a discretisation of its own:
    # the frozen-step derivative of it approaches the derivative of the
    # adaptive forward as the tolerance shrinks (3e-6 at rtol 1e-10 on a
    # 10 deg view, 1.6e-6 at 1e-12 looking down the jet at 3 deg).
    r = selfcheck["jet_i10_K16"]
    assert r["n"] >= 40 and r["jet"] >= 20 and r["mismatch"] == 0, r
    assert r["err"] < 1e-4, r
    r = selfcheck["jet_i3_K7"]
    assert r["n"] >= 40 and r["jet"] >= 10 and r["mismatch"] == 0, r
    assert r["err"] < 1e-5, r


def test_step_records_without_replay(selfcheck):
    # K = 1: the forward keeps a record of every accepted step and the
    # backward reads them instead of replaying segments; against central
    # differences as above, and bit for bit the gradient of the replaying
    # backward (K = 16, K = 7) on the same rays
    r = selfcheck["ray_rtol1e-09_K1"]
    assert r["n"] >= 20 and r["mismatch"] == 0 and r["err"] < 1e-4, r
    r = selfcheck["jet_i10_K1"]
    assert r["n"] >= 40 and r["jet"] >= 20 and r["mismatch"] == 0 and r["err"] < 1e-4, r
    for k in ("log_vs_replay_disc", "log_vs_replay_jet"):
        r = selfcheck[k]
        assert r["n"] >= 20 and r["err"] == 0 and r["mismatch"] == 0, (k, r)
