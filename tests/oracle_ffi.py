"""ctypes binding of oracle/liboracle.so -- the CPU restatement used as the
test oracle and as the CPU baseline.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")


class OrCamera(C.Structure):
    _fields_ = [("position", C.c_float * 4), ("view", C.c_float * 16), ("fov", C.c_float),
                ("_pad", C.c_float * 3)]


class OrWorld(C.Structure):
    _fields_ = [("kind", C.c_int), ("dim", C.c_uint32), ("grid", C.c_void_p),
                ("palette", C.c_void_p), ("chunks", C.c_void_p), ("pool", C.c_void_p),
                ("pool_words", C.c_uint64), ("atlas", C.c_void_p)]


class OrHit(C.Structure):
    _fields_ = [("data", C.c_uint32), ("hit_pos", C.c_float * 3), ("normal", C.c_float * 3),
                ("face", C.c_int), ("block", C.c_int32 * 3), ("sub", C.c_uint32),
                ("steps", C.c_int), ("tie_margin", C.c_float), ("lookups", C.c_uint32)]


class OrFrame(C.Structure):
    _fields_ = [("shaded", C.c_void_p), ("hits", C.c_void_p), ("color", C.c_void_p),
                ("normal", C.c_void_p), ("position", C.c_void_p), ("illumination", C.c_void_p),
                ("steps", C.c_void_p), ("tie_margin", C.c_void_p), ("stats", C.c_void_p),
                ("display", C.c_void_p)]


class OrPickResult(C.Structure):
    _fields_ = [("data", C.c_uint32), ("face", C.c_uint32), ("steps", C.c_uint32),
                ("_pad", C.c_uint32), ("hit_pos", C.c_float * 3), ("normal", C.c_float * 3),
                ("place", C.c_int32 * 3), ("remove", C.c_int32 * 3)]


class OrParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("max_steps_primary", C.c_int32), ("max_steps_shadow", C.c_int32),
                ("sun", C.c_float * 3), ("flags", C.c_uint32), ("traversal", C.c_uint32),
                ("ambient", C.c_float), ("albedo", C.c_void_p)]


_L = None


def lib():
    global _L
    if _L is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise OSError(f"{path} missing: run make -C oracle")
        L = C.CDLL(path)
        f3 = C.POINTER(C.c_float)
        L.or_tan_half_fov.restype = C.c_float
        L.or_tan_half_fov.argtypes = [C.c_float]
        L.or_raygen.argtypes = [C.POINTER(OrCamera), C.c_float, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_uint32, f3, f3]
        L.or_intersect_aabb.argtypes = [f3, f3, f3, f3, f3, f3]
        L.or_entry_origin.argtypes = [f3, f3, C.c_uint32, f3]
        L.or_trace_map.restype = C.c_int
        L.or_trace_map.argtypes = [C.POINTER(OrWorld), f3, f3, C.c_int, C.POINTER(OrHit),
                                   C.c_void_p, C.c_int]
        L.or_trace_entities.restype = C.c_int
        L.or_trace_entities.argtypes = [f3, f3, C.c_float]
        L.or_sky.argtypes = [f3, f3, f3]
        L.or_render.argtypes = [C.POINTER(OrWorld), C.POINTER(OrCamera), C.POINTER(OrParams),
                                C.POINTER(OrFrame), C.c_uint32, C.c_uint32, C.c_int]
        L.or_pick.argtypes = [C.POINTER(OrWorld), C.POINTER(OrCamera), C.c_int,
                              C.POINTER(OrPickResult)]
        L.or_present.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
        L.or_shade_lambert.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, f3, C.c_float,
                                       C.c_void_p]
        L.or_lambert_grad64.restype = C.c_double
        L.or_lambert_grad64.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, f3, C.c_float,
                                        C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_fnl_gradients.argtypes = [C.c_void_p]
        L.or_fnl_simplex2.restype = C.c_float
        L.or_fnl_simplex2.argtypes = [C.c_int, C.c_float, C.c_float]
        L.or_fnl_fbm2.restype = C.c_float
        L.or_fnl_fbm2.argtypes = [C.c_float, C.c_float]
        L.or_procgen_height.restype = C.c_uint32
        L.or_procgen_height.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_uint32,
                                        C.c_uint32]
        L.or_procgen.restype = C.c_int
        L.or_procgen.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_void_p, C.c_void_p,
                                 C.c_uint64, C.POINTER(C.c_uint64)]
        _L = L
    return _L


def procgen(dim, offset_x=0.0, offset_y=0.0):
    """procgen.zig:6-70 on an empty VoxelBrickmap(dim, 8): (status, chunks, pool)."""
    M = dim // 8
    cap = M * M * 8
    while True:
        chunks = np.zeros(M ** 3, np.uint32)
        pool = np.zeros(cap * 512, np.uint32)
        words = C.c_uint64()
        rc = lib().or_procgen(dim, offset_x, offset_y, chunks.ctypes.data, pool.ctypes.data,
                              pool.size, C.byref(words))
        if rc != 1:
            return rc, chunks, pool[:words.value].copy()
        cap *= 2


def procgen_heights(dim, offset_x=0.0, offset_y=0.0, xs=None, zs=None):
    """Column heights vh[x, z] (procgen.zig:22-23) of the given columns."""
    xs = range(dim) if xs is None else xs
    zs = range(dim) if zs is None else zs
    f = lib().or_procgen_height
    return np.array([[f(dim, offset_x, offset_y, x, z) for z in zs] for x in xs], np.uint32)


def fvec(v):
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


class DenseWorld:
    """Keeps the numpy buffers alive for the lifetime of the OrWorld."""

    def __init__(self, grid: np.ndarray, palette: np.ndarray):
        self.grid = np.ascontiguousarray(grid, dtype=np.uint8)
        self.palette = np.ascontiguousarray(palette, dtype=np.uint32)
        self.s = OrWorld(0, self.grid.shape[0], self.grid.ctypes.data, self.palette.ctypes.data,
                         None, None, 0, None)


class BrickmapWorld:
    def __init__(self, dim, chunks, pool, atlas):
        self.chunks = np.ascontiguousarray(chunks, dtype=np.uint32)
        self.pool = np.ascontiguousarray(pool, dtype=np.uint32)
        self.atlas = np.ascontiguousarray(atlas, dtype=np.uint32)
        self.s = OrWorld(1, dim, None, None, self.chunks.ctypes.data,
                         self.pool.ctypes.data if self.pool.size else None, self.pool.size,
                         self.atlas.ctypes.data)


def camera(cam) -> OrCamera:
    c = OrCamera()
    C.memmove(C.addressof(c), C.addressof(cam), 96)
    return c


def trace(world, o, d, max_steps, path_cap=0):
    h = OrHit()
    path = np.zeros((max(path_cap, 1), 3), dtype=np.int32)
    L = lib()
    L.or_trace_map(C.byref(world.s), fvec(o), fvec(d), int(max_steps), C.byref(h),
                   path.ctypes.data if path_cap else None, int(path_cap))
    return h, path[: min(h.steps, path_cap)] if path_cap else None


def cpu_threads():
    """Threads for the oracle: this process's CPU share, at most 16."""
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def render(world, cam, width, height, max_primary, max_shadow, sun, flags=0, traversal=0,
           ambient=0.0, albedo=None, rows=None, threads=None, gbuffer=True, display=False):
    """Full (or row-range) oracle render; returns a dict of numpy arrays."""
    threads = cpu_threads() if threads is None else threads
    p = OrParams(width, height, max_primary, max_shadow, fvec(sun), flags, traversal,
                 float(ambient), None if albedo is None else albedo.ctypes.data)
    out = {
        "shaded": np.zeros((height, width, 4), np.float32),
        "hits": np.zeros((height, width, 3), np.uint32),
        "steps": np.zeros((height, width), np.uint32),
        "tie_margin": np.zeros((height, width), np.float32),
        "stats": np.zeros(8, np.uint64),
    }
    if display:
        out["display"] = np.zeros((height, width), np.uint32)
    if gbuffer:
        for k in ("color", "normal", "position", "illumination"):
            out[k] = np.zeros((height, width, 4), np.float32)
    ptr = lambda k: out[k].ctypes.data if k in out else None
    f = OrFrame(ptr("shaded"), ptr("hits"), ptr("color"), ptr("normal"), ptr("position"),
                ptr("illumination"), ptr("steps"), ptr("tie_margin"), ptr("stats"),
                ptr("display"))
    y0, y1 = (0, height) if rows is None else rows
    oc = camera(cam)
    lib().or_render(C.byref(world.s), C.byref(oc), C.byref(p), C.byref(f), y0, y1, threads)
    return out


def present(composite):
    """Oracle present pass over an (H, W, 4) float32 composite: (out, display)."""
    c = np.ascontiguousarray(composite, dtype=np.float32)
    H, W = c.shape[:2]
    out = np.zeros_like(c)
    disp = np.zeros((H, W), np.uint32)
    lib().or_present(c.ctypes.data, W, H, out.ctypes.data, disp.ctypes.data)
    return out, disp


def pick(world, cam, max_steps=64) -> OrPickResult:
    r = OrPickResult()
    oc = camera(cam)
    lib().or_pick(C.byref(world.s), C.byref(oc), int(max_steps), C.byref(r))
    return r
