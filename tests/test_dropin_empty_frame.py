"""bh::render of a frame without pixels through the drop-in (CPU only).

The reference's render() (src/render.cpp:396-497) traces nothing for a 0 x H
or W x 0 frame: it returns the empty image, adds nothing to the counters and
assigns only the time.  blackhole_amd/integration/render_hip.cpp, compiled
against the reference's own render.h and linked with libbhx.so, must do the
same without touching a GPU -- there is none here, so any library call would
throw.  Skipped where the reference headers or the library are absent.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_SRC = "/root/reference/src"
LIB_DIR = os.path.join(ROOT, "blackhole_amd", "lib")

MAIN = r"""
#include "render.h"
#include <cstdio>
int main() {
    int bad = 0;
    const int shapes[3][2] = {{0, 48}, {64, 0}, {0, 0}};
    for (const auto& wh : shapes) {
        bh::RenderConfig cfg;
        cfg.quiet = true;
        cfg.width = wh[0];
        cfg.height = wh[1];
        bh::RenderStats st;
        st.rays = 7;
        st.max_norm_error = 0.5;
        st.seconds = -1.0;
        const img::Image im = bh::render(cfg, st);
        const bool ok = im.width == cfg.width && im.height == cfg.height && im.pixels.empty() &&
                        st.rays == 7 && st.max_norm_error == 0.5 && st.seconds >= 0.0 && st.seconds < 1.0;
        std::printf("%dx%d %s\n", cfg.width, cfg.height, ok ? "ok" : "BAD");
        bad += !ok;
    }
    return bad;
}
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "render.h")) or
                    not os.path.exists(os.path.join(LIB_DIR, "libbhx.so")),
                    reason="reference headers or libbhx.so absent")
def test_empty_frame_returns_empty_image_without_gpu(tmp_path):
    main = tmp_path / "main.cpp"
    main.write_text(MAIN)
    exe = tmp_path / "empty_frame"
    cmd = ["g++", "-std=c++17", "-O1", "-w", f"-I{REF_SRC}", f"-I{ROOT}/include", str(main),
           os.path.join(ROOT, "blackhole_amd", "integration", "render_hip.cpp"), "-o", str(exe), "-pthread",
           f"-L{LIB_DIR}", "-lbhx", f"-Wl,-rpath,{LIB_DIR}"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")   # no device, even where one exists
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, env=env)
    assert p.returncode == 0, p.stdout + p.stderr[-1000:]
    assert p.stdout.count(" ok") == 3, p.stdout
