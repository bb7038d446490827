"""Build the in-tree HIP library blackhole_amd/lib/libbhx.so for gfx950.

Plain hipcc, no JIT cache: the built .so lives in the tree so it travels with
the repository snapshot to the GPU host.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libbhx.so")

HIP_SOURCES = ["tracer.hip", "shade.hip", "api.hip", "order.hip", "batch.hip", "adjoint.hip", "display.hip",
               "png.hip", "noalloc.hip"]
# Per-file flags.  The tracer's loop uses ~70 distinct FP64 constants (DP5
# tableau, sincos/log/exp coefficients); FP64 VALU ops take no 64-bit
# literals, so machine LICM would pin every one of them in two registers for
# the whole loop and spill.  Without it they are rematerialised where used.
# tracer.hip is compiled with -ffp-contract=on as well: its kernels (single
# scene, lean, batched, pilot) inline one tracer body into differently shaped
# code, and with FMAs formed per source expression they round identically, so
# a ray's bits do not depend on which of them traced it.
# adjoint.hip: the recording forward and the backward's replay inline the same
# DP5 attempt into two kernels and must round identically, so FMAs are formed
# per source expression only (-ffp-contract=on), never across expressions by
# the backend, where the choice depends on the surrounding kernel.
# display.hip: the glare's sums must be the reference's IEEE sequence (no FMA).
# shade.hip: the same LICM setting as the tracer's (the shading loop inlines the
# library's acos / atan2 / exp / log / pow, whose constants would otherwise be
# hoisted into registers: 256 VGPRs and scratch spills instead of 144 and none),
# and per-expression FMAs like the tracer: the single-scene and the batched
# shading kernel inline one body and must round alike.
EXTRA = {"tracer.hip": ["-mllvm", "-disable-machine-licm", "-ffp-contract=on"],
         "noalloc.hip": ["-ffp-contract=off"],
         "shade.hip": ["-mllvm", "-disable-machine-licm", "-ffp-contract=on"],
         "adjoint.hip": ["-mllvm", "-disable-machine-licm", "-ffp-contract=on"],
         "display.hip": ["-ffp-contract=off"]}
CXX_SOURCES = ["host_setup.cpp", "host_physics.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BHX_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
# No packed FP32 instructions anywhere in the device code.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".hip", ".o"))
        _run([HIPCC, *COMMON, *DEVICE, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(LIBDIR, src.replace(".cpp", ".o"))
        _run(["g++", *COMMON, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=" + ARCH, *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
