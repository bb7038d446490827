"""The reference's program with the GPU render() linked in (oracle/Makefile
target `cli`) still passes the reference's own physics validation suite
(`--test`, src/validate.cpp), which runs on the host only.  Skipped where the
program was not built (the reference sources are needed to build it)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "blackhole_hip")


@pytest.mark.skipif(not os.path.exists(HIP_BIN), reason="reference program not built")
def test_validation_suite_passes_with_gpu_render_linked():
    p = subprocess.run([HIP_BIN, "--test"], capture_output=True, text=True, timeout=900)
    tail = (p.stdout + p.stderr)[-600:]
    assert p.returncode == 0, tail
    assert " 0 failed" in tail, tail
