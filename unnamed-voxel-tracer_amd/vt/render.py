"""Host-side handles over the C-ABI: contexts, worlds, frames, render calls.

Mirrors the reference's render boundary: ``Game.render`` (src/game.zig:232-256)
binds the camera UBO, the voxel brickmap SSBOs, the model atlas and the
G-buffer images, then dispatches the primary and shadow passes.  Here a
``World`` owns the device copy of the voxel data, a ``Frame`` owns the output
images (torch CUDA tensors are only the device-memory plumbing) and
``render`` enqueues the fused HIP launch on the current torch stream.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .scenes import SUN


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("vt needs a HIP device (torch.cuda.is_available() is False)")
    return torch


# ---- argument checks ---------------------------------------------------------
# The C-ABI takes raw device pointers and cannot see how large the buffers
# behind them are; these checks run before every call that launches work, so
# a mismatched tensor raises here instead of a kernel writing past it.

def _dtype_name(t):
    return str(getattr(t, "dtype", type(t).__name__)).replace("torch.", "")


def check_tensor(t, name, dtype, min_numel, device):
    """t must be a contiguous tensor on HIP device `device`, of `dtype` (a
    dtype name: "float32", "int32", "int64") with at least min_numel
    elements.  None passes (the C-ABI treats it as absent)."""
    if t is None:
        return
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{name}: needs a tensor on the HIP device (got {getattr(t, 'device', t)})")
    if t.device.index != device:
        raise ValueError(f"{name}: on cuda:{t.device.index}, the context is on cuda:{device}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if _dtype_name(t) != dtype:
        raise TypeError(f"{name}: dtype {_dtype_name(t)}, expected {dtype}")
    if t.numel() < min_numel:
        raise ValueError(f"{name}: {t.numel()} elements, needs at least {min_numel}")


def albedo_len(world, flags) -> int:
    """Floats an albedo (or its gradient) holds for this world and flags."""
    if flags & N.VT_FLAG_ALBEDO_COMPACT:
        if world.n_occupied is None:
            raise ValueError("VT_FLAG_ALBEDO_COMPACT: build the compact index first "
                             "(build_compact_index)")
        return 3 * world.n_occupied
    return 3 * world.dim ** 3


def check_params(ctx, world, prm, who="render"):
    """The device buffers params() recorded, against the world and flags."""
    keep = getattr(prm, "_tensors", None)
    if keep is None:
        if prm.albedo or prm.grad_albedo or prm.grad_sun or prm.loss:
            raise ValueError(f"{who}: RenderParams with device buffers must come from vt.params()")
        return
    dev = ctx.device
    lam = bool(prm.flags & N.VT_FLAG_LAMBERT)
    if lam or keep["albedo"] is not None or keep["grad_albedo"] is not None:
        n = albedo_len(world, prm.flags) if world.kind == "dense" else 0
        check_tensor(keep["albedo"], f"{who}: albedo", "float32", n, dev)
        check_tensor(keep["grad_albedo"], f"{who}: grad_albedo", "float32", n, dev)
    check_tensor(keep["grad_sun"], f"{who}: grad_sun", "float32", 3, dev)
    check_tensor(keep["loss"], f"{who}: loss", "float32", 1, dev)


class Context:
    def __init__(self, device: int = 0):
        self.device = device
        h = C.c_void_p()
        N.check(N.lib().vt_create(C.byref(h), device), "vt_create")
        self.h = h

    def info(self):
        cu = C.c_int()
        name = C.create_string_buffer(256)
        N.check(N.lib().vt_device_info(self.h, C.byref(cu), name, 256), "vt_device_info")
        return {"compute_units": cu.value, "arch": name.value.decode()}

    def close(self):
        if self.h:
            N.check(N.lib().vt_destroy(self.h), "vt_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_brickmap_arrays(dim, chunks, pool, atlas):
    """Sizes vt_world_create_brickmap reads from the host arrays."""
    if dim < 8 or dim % 8:
        raise ValueError(f"brickmap: dim {dim} must be a positive multiple of 8")
    if chunks.size != (dim // 8) ** 3:
        raise ValueError(f"brickmap: {chunks.size} chunk entries, a {dim}^3 map has "
                         f"{(dim // 8) ** 3}")
    if pool.size % 512:
        raise ValueError(f"brickmap: pool of {pool.size} words is not whole 8^3 blocks")
    if atlas.size != 256 ** 3:
        raise ValueError(f"brickmap: atlas of {atlas.size} texels, expected 256^3")


class World:
    def __init__(self, ctx: Context, h: C.c_void_p, dim: int, kind: str):
        self.ctx, self.h, self.dim, self.kind = ctx, h, dim, kind
        self.n_occupied = None  # compact index size, once built

    @classmethod
    def dense(cls, ctx: Context, grid: np.ndarray, palette: np.ndarray) -> "World":
        grid = np.ascontiguousarray(grid, dtype=np.uint8)
        dim = grid.shape[0]
        assert grid.shape == (dim, dim, dim)
        pal = np.ascontiguousarray(palette, dtype=np.uint32)
        assert pal.shape == (256,)
        h = C.c_void_p()
        N.check(N.lib().vt_world_create_dense_u8(ctx.h, grid.ctypes.data, dim, pal.ctypes.data,
                                                 C.byref(h)), "vt_world_create_dense_u8")
        return cls(ctx, h, dim, "dense")

    @classmethod
    def brickmap(cls, ctx: Context, dim: int, chunks: np.ndarray, pool: np.ndarray,
                 atlas: np.ndarray) -> "World":
        chunks = np.ascontiguousarray(chunks, dtype=np.uint32).reshape(-1)
        pool = np.ascontiguousarray(pool, dtype=np.uint32).reshape(-1)
        atlas = np.ascontiguousarray(atlas, dtype=np.uint32).reshape(-1)
        check_brickmap_arrays(dim, chunks, pool, atlas)
        h = C.c_void_p()
        N.check(N.lib().vt_world_create_brickmap(ctx.h, dim, chunks.ctypes.data,
                                                 pool.ctypes.data if pool.size else None,
                                                 pool.size, atlas.ctypes.data, C.byref(h)),
                "vt_world_create_brickmap")
        return cls(ctx, h, dim, "brickmap")

    @classmethod
    def empty_brickmap(cls, ctx: Context, dim: int, atlas: np.ndarray) -> "World":
        """VoxelBrickmap.init (voxel.zig:32-39): every chunk empty, no pool blocks."""
        M = dim // 8
        return cls.brickmap(ctx, dim, np.zeros(M ** 3, np.uint32), np.zeros(0, np.uint32), atlas)

    def set_voxels(self, xyz, values, stream=None):
        """Edits in order (last write of a voxel wins); see vt_world_set_voxels."""
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32).reshape(-1, 3)
        values = np.ascontiguousarray(values, dtype=np.uint32).reshape(-1)
        assert len(values) == len(xyz)
        N.check(N.lib().vt_world_set_voxels(self.ctx.h, self.h, xyz.ctypes.data,
                                            values.ctypes.data, len(values),
                                            C.c_void_p(_stream(stream))), "vt_world_set_voxels")
        if self.n_occupied is not None:  # edits renumber the compact index
            build_compact_index(self.ctx, self)

    def get_voxels(self, xyz, stream=None) -> np.ndarray:
        xyz = np.ascontiguousarray(xyz, dtype=np.uint32).reshape(-1, 3)
        out = np.zeros(len(xyz), np.uint32)
        N.check(N.lib().vt_world_get_voxels(self.ctx.h, self.h, xyz.ctypes.data, len(xyz),
                                            out.ctypes.data, C.c_void_p(_stream(stream))),
                "vt_world_get_voxels")
        return out

    def read_brickmap(self, stream=None):
        """(chunks, pool) of a brickmap world, as VoxelBrickmap's mappings hold them."""
        M = self.dim // 8
        words = C.c_size_t()
        N.check(N.lib().vt_world_read_brickmap(self.ctx.h, self.h, None, None, 0, C.byref(words),
                                               C.c_void_p(_stream(stream))),
                "vt_world_read_brickmap")
        chunks = np.zeros(M ** 3, np.uint32)
        pool = np.zeros(words.value, np.uint32)
        N.check(N.lib().vt_world_read_brickmap(self.ctx.h, self.h, chunks.ctypes.data,
                                               pool.ctypes.data if pool.size else None,
                                               pool.size, C.byref(words),
                                               C.c_void_p(_stream(stream))),
                "vt_world_read_brickmap")
        return chunks, pool

    def read_free_map(self, stream=None):
        """(bitmaps, valid): the free-region map as a uint32 array [8, words per
        octant] (octant o = x negative | y negative << 1 | z negative << 2; bit r
        = region with Morton code r free), None for a world without one."""
        wpo, valid = C.c_uint32(), C.c_int32()
        N.check(N.lib().vt_world_read_free_map(self.ctx.h, self.h, None, 0, C.byref(wpo),
                                               C.byref(valid), C.c_void_p(_stream(stream))),
                "vt_world_read_free_map")
        words = np.zeros(8 * wpo.value, np.uint32)
        N.check(N.lib().vt_world_read_free_map(self.ctx.h, self.h, words.ctypes.data, words.size,
                                               C.byref(wpo), C.byref(valid),
                                               C.c_void_p(_stream(stream))),
                "vt_world_read_free_map")
        return words.reshape(8, wpo.value), bool(valid.value)

    def procgen(self, offset_x: float = 0.0, offset_y: float = 0.0, stream=None):
        """procgen.zig:6-70 on this (empty) brickmap world; see vt_world_procgen."""
        N.check(N.lib().vt_world_procgen(self.ctx.h, self.h, offset_x, offset_y,
                                         C.c_void_p(_stream(stream))), "vt_world_procgen")

    def pick(self, cam: N.Camera, max_steps: int = 64, stream=None) -> N.PickResult:
        """Centre-ray pick of terrain_edit.comp.glsl (64 steps there)."""
        r = N.PickResult()
        N.check(N.lib().vt_pick(self.ctx.h, self.h, C.byref(cam), max_steps, C.byref(r),
                                C.c_void_p(_stream(stream))), "vt_pick")
        return r

    def device_bytes(self) -> int:
        d, b = C.c_uint32(), C.c_size_t()
        N.check(N.lib().vt_world_info(self.h, C.byref(d), C.byref(b)), "vt_world_info")
        return b.value

    def close(self):
        if self.h:
            N.check(N.lib().vt_world_destroy(self.h), "vt_world_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Frame:
    """Device output images of one view (torch tensors on the context's GPU)."""
    width: int
    height: int
    shaded: object = None
    hits: object = None
    color: object = None
    normal: object = None
    position: object = None
    illumination: object = None
    steps: object = None
    stats: object = None
    target: object = None
    display: object = None
    schedule: object = None

    @classmethod
    def alloc(cls, width, height, device=0, shaded=True, hits=True, gbuffer=False, steps=False,
              stats=False, display=False, schedule=True):
        """Device buffers for a width x height view.  `schedule`: a zeroed
        tile schedule buffer (vt_frame.schedule), so that each render into
        this frame launches its blocks in the cost order the previous render
        measured."""
        torch = _torch()
        dev = f"cuda:{device}"
        f = cls(width, height)
        img = lambda: torch.empty((height, width, 4), dtype=torch.float32, device=dev)
        if shaded:
            f.shaded = img()
        if hits:
            f.hits = torch.empty((height, width, 3), dtype=torch.int32, device=dev)
        if gbuffer:
            f.color, f.normal, f.position, f.illumination = img(), img(), img(), img()
        if steps:
            f.steps = torch.empty((height, width), dtype=torch.int32, device=dev)
        if stats:
            f.stats = torch.zeros(8, dtype=torch.int64, device=dev)
        if display:
            f.display = torch.empty((height, width), dtype=torch.int32, device=dev)
        if schedule:
            f.schedule = torch.zeros(N.schedule_words(width, height), dtype=torch.int32,
                                     device=dev)
        return f

    def check(self, device, who="frame"):
        """Every buffer large enough for width x height, on `device`."""
        n = self.width * self.height
        for name in ("shaded", "color", "normal", "position", "illumination", "target"):
            check_tensor(getattr(self, name), f"{who}.{name}", "float32", 4 * n, device)
        check_tensor(self.hits, f"{who}.hits", "int32", 3 * n, device)
        check_tensor(self.steps, f"{who}.steps", "int32", n, device)
        check_tensor(self.display, f"{who}.display", "int32", n, device)
        check_tensor(self.stats, f"{who}.stats", "int64", 8, device)
        check_tensor(self.schedule, f"{who}.schedule", "int32",
                     N.schedule_words(self.width, self.height), device)

    def struct(self) -> N.Frame:
        p = lambda t: None if t is None else t.data_ptr()
        return N.Frame(p(self.shaded), p(self.hits), p(self.color), p(self.normal),
                       p(self.position), p(self.illumination), p(self.steps), p(self.stats),
                       p(self.target), p(self.display), p(self.schedule))

    def stats_dict(self):
        s = self.stats.cpu().numpy()
        return {k: int(v) for k, v in zip(N.STAT_NAMES, s)}


def params(width, height, max_primary=512, max_shadow=512, sun=SUN, flags=0, traversal=0,
           ambient=0.0, albedo=None, grad_albedo=None, grad_sun=None, loss=None
           ) -> N.RenderParams:
    p = N.RenderParams()
    p.width, p.height = width, height
    p.max_steps_primary, p.max_steps_shadow = max_primary, max_shadow
    p.sun[:] = [float(np.float32(x)) for x in sun]
    p.flags, p.traversal, p.ambient = flags, traversal, float(ambient)
    ptr = lambda t: None if t is None else t.data_ptr()
    p.albedo, p.grad_albedo, p.grad_sun, p.loss = (ptr(albedo), ptr(grad_albedo), ptr(grad_sun),
                                                   ptr(loss))
    # the tensors behind the pointers: kept alive and checked before launches
    p._tensors = dict(albedo=albedo, grad_albedo=grad_albedo, grad_sun=grad_sun, loss=loss)
    return p


def present(ctx: Context, composite, out=None, display=None, stream=None):
    """Present pass (crosshair + vignette) over an (H, W, 4) float32 composite."""
    H, W = composite.shape[0], composite.shape[1]
    check_tensor(composite, "present: composite", "float32", 4 * W * H, ctx.device)
    check_tensor(out, "present: out", "float32", 4 * W * H, ctx.device)
    check_tensor(display, "present: display", "int32", W * H, ctx.device)
    s = current_stream() if stream is None else stream
    ptr = lambda t: None if t is None else t.data_ptr()
    N.check(N.lib().vt_present(ctx.h, composite.data_ptr(), W, H, ptr(out), ptr(display),
                               C.c_void_p(s)), "vt_present")


def build_compact_index(ctx: Context, world: World) -> int:
    n = C.c_uint64()
    N.check(N.lib().vt_world_build_compact_index(ctx.h, world.h, C.byref(n)),
            "vt_world_build_compact_index")
    world.n_occupied = n.value
    return n.value


def _records(ctx, hits, who) -> int:
    """Hit records: an int32 [..., 3] tensor; returns their count."""
    if hits is None:
        raise ValueError(f"{who}: hits required")
    if hits.numel() % 3:
        raise ValueError(f"{who}: hits has {hits.numel()} words, not whole 3-word records")
    check_tensor(hits, f"{who}: hits", "int32", 3, ctx.device)
    return hits.numel() // 3


def lambert_shade(ctx: Context, world: World, hits, prm: N.RenderParams, out, stream=None):
    """out[i] = Lambert shading of hit record i (hits: int32 [..., 3] tensor)."""
    npix = _records(ctx, hits, "lambert_shade")
    check_params(ctx, world, prm, "lambert_shade")
    check_tensor(out, "lambert_shade: out", "float32", 4 * npix, ctx.device)
    s = current_stream() if stream is None else stream
    N.check(N.lib().vt_lambert_shade(ctx.h, world.h, hits.data_ptr(), hits.numel() // 3,
                                     C.byref(prm), out.data_ptr(), C.c_void_p(s)),
            "vt_lambert_shade")
    return out


def lambert_backward(ctx: Context, world: World, hits, prm: N.RenderParams, upstream, kind,
                     grad_albedo, grad_sun=None, loss=None, stream=None):
    npix = _records(ctx, hits, "lambert_backward")
    check_params(ctx, world, prm, "lambert_backward")
    check_tensor(upstream, "lambert_backward: upstream", "float32", 4 * npix, ctx.device)
    check_tensor(grad_albedo, "lambert_backward: grad_albedo", "float32",
                 albedo_len(world, prm.flags), ctx.device)
    check_tensor(grad_sun, "lambert_backward: grad_sun", "float32", 3, ctx.device)
    check_tensor(loss, "lambert_backward: loss", "float32", 1, ctx.device)
    s = current_stream() if stream is None else stream
    ptr = lambda t: None if t is None else t.data_ptr()
    N.check(N.lib().vt_lambert_backward(ctx.h, world.h, hits.data_ptr(), hits.numel() // 3,
                                        C.byref(prm), upstream.data_ptr(), kind,
                                        grad_albedo.data_ptr(), ptr(grad_sun), ptr(loss),
                                        C.c_void_p(s)), "vt_lambert_backward")


def tuning(tile=N.VT_TILE_8X8, chunk=1, traversal=N.VT_TRAVERSAL_FLAT, blocks_per_cu=8):
    return N.Tuning(tile, chunk, traversal, blocks_per_cu)


class Batch:
    """Many views in one persistent launch (vt_batch_*); `views` is a list of
    (camera, Frame) pairs whose frames stay alive as long as the batch."""

    def __init__(self, ctx: Context, world: World, views, prm: N.RenderParams,
                 tune: N.Tuning, stats=None):
        self.ctx, self.world, self.frames = ctx, world, [f for _, f in views]
        _check_views(ctx, world, views, prm, "Batch")
        check_tensor(stats, "Batch: stats", "int64", 8, ctx.device)
        self.prm = prm  # keeps the albedo / gradient tensors alive
        self.table = (N.View * len(views))()
        for i, (cam, f) in enumerate(views):
            self.table[i].camera = cam
            self.table[i].width, self.table[i].height = f.width, f.height
            self.table[i].frame = f.struct()
        self.stats = stats
        h = C.c_void_p()
        N.check(N.lib().vt_batch_create(ctx.h, world.h, self.table, len(views), C.byref(prm),
                                        C.byref(tune),
                                        None if stats is None else stats.data_ptr(),
                                        C.byref(h)), "vt_batch_create")
        self.h = h

    def launch(self, stream=None):
        s = current_stream() if stream is None else stream
        N.check(N.lib().vt_batch_launch(self.h, C.c_void_p(s)), "vt_batch_launch")

    def capture_graph(self):
        N.check(N.lib().vt_batch_capture_graph(self.h), "vt_batch_capture_graph")

    def graph_launch(self, stream=None):
        s = current_stream() if stream is None else stream
        N.check(N.lib().vt_batch_graph_launch(self.h, C.c_void_p(s)), "vt_batch_graph_launch")

    def close(self):
        if self.h:
            N.check(N.lib().vt_batch_destroy(self.h), "vt_batch_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_views(ctx, world, views, prm, who):
    check_params(ctx, world, prm, who)
    for i, (_, f) in enumerate(views):
        f.check(ctx.device, f"{who}: view {i}")


def autotune(ctx: Context, world: World, views, prm: N.RenderParams, reps=3):
    _check_views(ctx, world, views, prm, "autotune")
    table = (N.View * len(views))()
    for i, (cam, f) in enumerate(views):
        table[i].camera = cam
        table[i].width, table[i].height = f.width, f.height
        table[i].frame = f.struct()
    best, ms = N.Tuning(), C.c_float()
    N.check(N.lib().vt_batch_autotune(ctx.h, world.h, table, len(views), C.byref(prm), reps,
                                      C.byref(best), C.byref(ms)), "vt_batch_autotune")
    return best, ms.value


def current_stream():
    return _torch().cuda.current_stream().cuda_stream


def _stream(stream):
    return current_stream() if stream is None else stream


def free_map_used(world: World, frame: Frame, prm: N.RenderParams) -> bool:
    """Whether render() with these arguments ends rays early in free regions
    (vt_free_map_used)."""
    used = C.c_int32(0)
    N.check(N.lib().vt_free_map_used(world.h, C.byref(prm), C.byref(frame.struct()),
                                     C.byref(used)), "vt_free_map_used")
    return bool(used.value)


def render(ctx: Context, world: World, cam: N.Camera, frame: Frame, prm: N.RenderParams,
           stream=None) -> Frame:
    if (prm.width, prm.height) != (frame.width, frame.height):
        raise ValueError(f"render: params are {prm.width}x{prm.height}, the frame "
                         f"{frame.width}x{frame.height}")
    frame.check(ctx.device, "render: frame")
    check_params(ctx, world, prm, "render")
    fs = frame.struct()
    s = current_stream() if stream is None else stream
    N.check(N.lib().vt_render(ctx.h, world.h, C.byref(cam), C.byref(prm), C.byref(fs),
                              C.c_void_p(s)), "vt_render")
    return frame
