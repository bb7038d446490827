"""The C ABI: libbhx.so loads without a GPU, exports every symbol that
include/bhx.h declares, and the ctypes mirror has the C struct layouts."""
import ctypes as C
import os
import re
import subprocess

import pytest

from blackhole_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bhx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bhx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    names = declared_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.SYMBOLS) <= set(names)
    assert lib.bhx_abi_version() == 3


def test_no_gpu_call_needed_for_errors():
    lib = native.lib()
    # NULL context / arguments are rejected before any HIP call
    assert lib.bhx_create(0, None) == -1
    assert lib.bhx_render(None, None, None, None) == -1
    assert lib.bhx_destroy(None) == 0


def test_noalloc_trace_refuses_frames_beyond_its_limit():
    """bhx_trace_noalloc_dev holds at most 2^20 samples and 2^15 tiles of
    8 x 8 pixels in its static scratch array; larger frames are refused with
    BHX_E_TOO_LARGE and a message before any HIP call."""
    from blackhole_amd import scene as S
    lib = native.lib()
    rays = abi.Rays()
    for w, h, spp in ((1025, 1024, 1), (512, 512, 3), ((1 << 18) + 8, 2, 1)):
        sc = S.plain_frame(w, h)
        sc.sqrt_spp = spp
        assert lib.bhx_trace_noalloc_dev(0, C.byref(sc), C.byref(rays), None, None) == -4, (w, h, spp)
        assert b"exceeds" in lib.bhx_last_error(None)
    assert lib.bhx_trace_noalloc_dev(0, None, C.byref(rays), None, None) == -1
    assert lib.bhx_trace_noalloc_dev(-1, C.byref(S.plain_frame(8, 8)), C.byref(rays), None, None) == -1
    assert lib.bhx_last_error(None) == b"bhx_trace_noalloc_dev: device index out of range"


@pytest.mark.parametrize("struct,cname", [(abi.Scene, "bhx_scene"), (abi.Stats, "bhx_stats"),
                                          (abi.Rays, "bhx_rays"), (abi.Shaded, "bhx_shaded"),
                                          (abi.Launch, "bhx_launch"), (abi.Bytes, "bhx_bytes")])
def test_struct_layout_matches_header(tmp_path, struct, cname):
    fields = [f[0] for f in struct._fields_]
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           f'printf("%zu\\n", sizeof({cname}));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    assert vals[0] == C.sizeof(struct)
    for f, off in zip(fields, vals[1:]):
        assert getattr(struct, f).offset == off, f
