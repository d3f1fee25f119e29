"""ctypes binding of libvt.so (include/vt.h).

The library is loaded from ``unnamed-voxel-tracer_amd/lib``; a missing
library raises immediately -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(_PKG, "lib")

VT_OK = 0
STATUS = {
    0: "VT_OK",
    1: "VT_ERR_INVALID_ARGUMENT",
    2: "VT_ERR_NULL_POINTER",
    3: "VT_ERR_OUT_OF_MEMORY",
    4: "VT_ERR_DEVICE",
    5: "VT_ERR_NO_DEVICE",
    6: "VT_ERR_UNSUPPORTED",
    7: "VT_ERR_MISALIGNED",
    8: "VT_ERR_WRONG_CONTEXT",
}
VT_ERR_INVALID_ARGUMENT = 1
VT_ERR_NULL_POINTER = 2
VT_ERR_OUT_OF_MEMORY = 3
VT_ERR_DEVICE = 4
VT_ERR_NO_DEVICE = 5
VT_ERR_UNSUPPORTED = 6
VT_ERR_MISALIGNED = 7
VT_ERR_WRONG_CONTEXT = 8

VT_FLAG_GBUF_UNORM8 = 1 << 0
VT_FLAG_LAMBERT = 1 << 1
VT_FLAG_ALBEDO_COMPACT = 1 << 2
VT_FLAG_GRAD_L2 = 1 << 3
VT_FLAG_PRESENT = 1 << 4
VT_FLAG_FREE_MAP_ON = 1 << 5
VT_FLAG_FREE_MAP_OFF = 1 << 6
VT_FREE_MIN_CAP = 512
VT_FREE_MIN_SHARE = 16
VT_UPSTREAM_GRADIENT = 0
VT_UPSTREAM_L2_TARGET = 1
VT_TRAVERSAL_FLAT = 0
VT_TRAVERSAL_BITMASK = 1
VT_TRAVERSAL_BITMASK8 = 2

STAT_NAMES = ["primary_steps", "shadow_steps", "primary_hits", "occluded",
              "l1_fetches", "l0_fetches", "leaf_reads", "entity_hits"]


class Camera(C.Structure):
    """Byte layout of Camera.UniformData (camera.zig:12-16)."""
    _fields_ = [("position", C.c_float * 4), ("view", C.c_float * 16),
                ("fov", C.c_float), ("_pad", C.c_float * 3)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("max_steps_primary", C.c_int32), ("max_steps_shadow", C.c_int32),
                ("sun", C.c_float * 3), ("flags", C.c_uint32), ("traversal", C.c_uint32),
                ("ambient", C.c_float), ("albedo", C.c_void_p), ("grad_albedo", C.c_void_p),
                ("grad_sun", C.c_void_p), ("loss", C.c_void_p)]


class Frame(C.Structure):
    _fields_ = [("shaded", C.c_void_p), ("hits", C.c_void_p), ("color", C.c_void_p),
                ("normal", C.c_void_p), ("position", C.c_void_p),
                ("illumination", C.c_void_p), ("steps", C.c_void_p), ("stats", C.c_void_p),
                ("target", C.c_void_p), ("display", C.c_void_p), ("schedule", C.c_void_p)]


class View(C.Structure):
    _fields_ = [("camera", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("_pad", C.c_uint32 * 2), ("frame", Frame)]


class Tuning(C.Structure):
    _fields_ = [("tile", C.c_uint32), ("chunk", C.c_uint32), ("traversal", C.c_uint32),
                ("blocks_per_cu", C.c_uint32)]


class PickResult(C.Structure):
    _fields_ = [("data", C.c_uint32), ("face", C.c_uint32), ("steps", C.c_uint32),
                ("_pad", C.c_uint32), ("hit_pos", C.c_float * 3), ("normal", C.c_float * 3),
                ("place", C.c_int32 * 3), ("remove", C.c_int32 * 3)]


VT_TILE_8X8, VT_TILE_16X4, VT_TILE_32X2 = 3, 4, 5
VT_NOALLOC_MAX_VIEWS = 256


def schedule_blocks(width: int, height: int) -> int:
    """VT_SCHEDULE_BLOCKS(W, H) of include/vt.h."""
    return ((width + 15) // 16) * ((height + 7) // 8)


def schedule_words(width: int, height: int) -> int:
    """VT_SCHEDULE_WORDS(W, H) of include/vt.h: words of a frame's tile
    schedule buffer."""
    return 512 + 3 * (4096 + 128 * ((schedule_blocks(width, height) + 7) // 8))

assert C.sizeof(Camera) == 96
assert C.sizeof(RenderParams) == 72
assert C.sizeof(Frame) == 88
assert C.sizeof(View) == 200
assert C.sizeof(PickResult) == 64


class VtError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)}: {detail}")


_lib = None


def lib() -> C.CDLL:
    """Load libvt.so (raises OSError when it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(LIB_DIR, "libvt.so")
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with unnamed-voxel-tracer_amd/build.py")
    L = C.CDLL(path)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        "vt_create": (i32, [C.POINTER(vp), i32]),
        "vt_destroy": (i32, [vp]),
        "vt_sync": (i32, [vp, vp]),
        "vt_status_string": (C.c_char_p, [i32]),
        "vt_last_error": (C.c_char_p, []),
        "vt_device_info": (i32, [vp, C.POINTER(i32), C.c_char_p, C.c_size_t]),
        "vt_world_create_dense_u8": (i32, [vp, vp, u32, vp, C.POINTER(vp)]),
        "vt_world_create_brickmap": (i32, [vp, u32, vp, vp, C.c_size_t, vp, C.POINTER(vp)]),
        "vt_world_destroy": (i32, [vp]),
        "vt_world_info": (i32, [vp, C.POINTER(u32), C.POINTER(C.c_size_t)]),
        "vt_render": (i32, [vp, vp, C.POINTER(Camera), C.POINTER(RenderParams),
                            C.POINTER(Frame), vp]),
        "vt_free_map_used": (i32, [vp, C.POINTER(RenderParams), C.POINTER(Frame),
                                   C.POINTER(i32)]),
        "vt_world_read_free_map": (i32, [vp, vp, vp, C.c_size_t, C.POINTER(u32), C.POINTER(i32),
                                         vp]),
        "vt_present": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "vt_vox_info": (i32, [vp, C.c_size_t, C.POINTER(u32), C.POINTER(u32)]),
        "vt_atlas_load_vox": (i32, [vp, C.c_size_t, vp, C.POINTER(u32)]),
        "vt_world_set_voxels": (i32, [vp, vp, vp, vp, u32, vp]),
        "vt_world_get_voxels": (i32, [vp, vp, vp, u32, vp, vp]),
        "vt_world_read_brickmap": (i32, [vp, vp, vp, vp, C.c_size_t, C.POINTER(C.c_size_t), vp]),
        "vt_world_procgen": (i32, [vp, vp, C.c_float, C.c_float, vp]),
        "vt_pick": (i32, [vp, vp, C.POINTER(Camera), i32, C.POINTER(PickResult), vp]),
        "vt_world_build_compact_index": (i32, [vp, vp, C.POINTER(C.c_uint64)]),
        "vt_lambert_shade": (i32, [vp, vp, vp, u32, C.POINTER(RenderParams), vp, vp]),
        "vt_lambert_backward": (i32, [vp, vp, vp, u32, C.POINTER(RenderParams), vp, u32, vp, vp,
                                      vp, vp]),
        "vt_batch_create": (i32, [vp, vp, C.POINTER(View), u32, C.POINTER(RenderParams),
                                  C.POINTER(Tuning), vp, C.POINTER(vp)]),
        "vt_batch_launch": (i32, [vp, vp]),
        "vt_batch_capture_graph": (i32, [vp]),
        "vt_batch_graph_launch": (i32, [vp, vp]),
        "vt_batch_destroy": (i32, [vp]),
        "vt_batch_autotune": (i32, [vp, vp, C.POINTER(View), u32, C.POINTER(RenderParams),
                                    i32, C.POINTER(Tuning), C.POINTER(C.c_float)]),
        "vt_batch_render_noalloc": (i32, [vp, vp, C.POINTER(View), u32,
                                          C.POINTER(RenderParams), C.POINTER(Tuning), vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int, where: str) -> None:
    if status != VT_OK:
        detail = lib().vt_last_error().decode(errors="replace")
        raise VtError(status, where, detail)


def exported_symbols() -> list[str]:
    """Function names declared in include/vt.h."""
    import re
    hdr = os.path.join(os.path.dirname(_PKG), "include", "vt.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(vt_\w+)\s*\(", text, re.M)))
