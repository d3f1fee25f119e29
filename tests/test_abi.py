"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/vt.h declares, validates arguments before touching the device,
and its structs have the layout the Python mirror and the reference's
Camera.UniformData (camera.zig:12-16) assume."""
import ctypes as C
import os
import subprocess
import sys

import pytest

import vt
from vt import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = vt.lib()
    names = N.exported_symbols()
    assert {"vt_create", "vt_render", "vt_world_create_dense_u8",
            "vt_world_create_brickmap"} <= set(names)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(N.LIB_DIR, "libvt.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_product_library_does_not_link_the_oracle():
    out = subprocess.run(["ldd", os.path.join(N.LIB_DIR, "libvt.so")], capture_output=True,
                         text=True).stdout
    assert "oracle" not in out
    pkg = os.path.join(ROOT, "unnamed-voxel-tracer_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".c", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle_ffi" not in text and "liboracle" not in text, f


def test_status_strings():
    L = vt.lib()
    seen = set()
    for code in range(0, 9):
        s = L.vt_status_string(code).decode()
        assert s and s != "unknown status"
        seen.add(s)
    assert len(seen) == 9
    assert L.vt_status_string(99).decode() == "unknown status"


def test_null_arguments_rejected_before_device_work():
    L = vt.lib()
    assert L.vt_create(None, 0) == N.VT_ERR_NULL_POINTER
    assert L.vt_destroy(None) == N.VT_ERR_NULL_POINTER
    assert L.vt_world_destroy(None) == N.VT_ERR_NULL_POINTER
    out = C.c_void_p()
    assert L.vt_world_create_dense_u8(None, None, 16, None, C.byref(out)) == N.VT_ERR_NULL_POINTER
    assert out.value is None
    assert L.vt_world_create_dense_u8(None, None, 16, None, None) == N.VT_ERR_NULL_POINTER
    assert L.vt_world_create_brickmap(None, 8, None, None, 0, None, C.byref(out)) == \
        N.VT_ERR_NULL_POINTER
    assert L.vt_render(None, None, None, None, None, None) == N.VT_ERR_NULL_POINTER
    assert "vt_render" in L.vt_last_error().decode()
    assert L.vt_sync(None, None) == N.VT_ERR_NULL_POINTER
    assert L.vt_world_info(None, None, None) == N.VT_ERR_NULL_POINTER


def test_no_device_reported_without_gpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    h = C.c_void_p()
    assert vt.lib().vt_create(C.byref(h), 0) == N.VT_ERR_NO_DEVICE
    assert h.value is None


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "vt.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(vt_camera), offsetof(vt_camera, view),
         offsetof(vt_camera, fov), sizeof(vt_hit_record));
  printf("%zu %zu %zu %zu\n", sizeof(vt_render_params), offsetof(vt_render_params, sun),
         offsetof(vt_render_params, albedo), sizeof(vt_frame));
  printf("%zu %zu %zu\n", offsetof(vt_frame, schedule), sizeof(vt_view), offsetof(vt_view, frame));
  const unsigned wh[][2] = {{1, 1}, {256, 256}, {1280, 720}, {1920, 1080}, {17, 1001}, {16384, 16384}};
  for (int i = 0; i < 6; i++)
    printf("%u %u %u ", (unsigned)VT_SCHEDULE_BLOCKS(wh[i][0], wh[i][1]),
           (unsigned)VT_SCHEDULE_WORDS(wh[i][0], wh[i][1]), wh[i][0] * 100000u + wh[i][1]);
  printf("\n");
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    a, b, c, d = subprocess.run([str(exe)], capture_output=True, text=True,
                                check=True).stdout.split("\n")[:4]
    # std140 u_Camera: vec4 @0, mat4 @16, float @80, 96 bytes (camera.glsl:2-6)
    assert [int(x) for x in a.split()] == [96, 16, 80, 12]
    assert [int(x) for x in b.split()] == [
        C.sizeof(N.RenderParams), N.RenderParams.sun.offset, N.RenderParams.albedo.offset,
        C.sizeof(N.Frame)]
    assert N.Camera.view.offset == 16 and N.Camera.fov.offset == 80
    assert [int(x) for x in c.split()] == [N.Frame.schedule.offset, C.sizeof(N.View),
                                           N.View.frame.offset]
    vals = [int(x) for x in d.split()]
    for i in range(0, len(vals), 3):
        W, H = divmod(vals[i + 2], 100000)
        assert vals[i] == N.schedule_blocks(W, H) and vals[i + 1] == N.schedule_words(W, H)
        # the blocks of either tile shape vt_render uses fit the lists
        assert N.schedule_blocks(W, H) >= max(-(-W // 32) * -(-H // 8), -(-W // 16) * -(-H // 16))


def test_library_is_the_build_of_these_sources():
    """The build record beside libvt.so names this library and these sources
    (rebuild after changing a source)."""
    import build as B
    p = B.provenance()
    assert p["record"] is not None, "no build record: run unnamed-voxel-tracer_amd/build.py"
    assert p["lib_sha256_matches"], "libvt.so differs from the build that wrote the record"
    assert p["sources_sha256_matches"], "sources changed since libvt.so was built"
