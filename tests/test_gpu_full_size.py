"""The 1024^3 configurations at their full size, on the GPU.

* C2 world (1024^3, terrain x8 + 32,000 sprites), its 64 uneven cameras:
  bands of rows of several views against the oracle (both kernel variants,
  bit-exact hit records, step counts and G-buffer); every view's batch
  output bit-identical to a plain render of that view; the flat and bitmask
  traversals identical on every view.
* C4 (the same world, 48 views, compact albedo index): the Lambert batch
  forward bit-identical to plain per-view renders, and the compact index a
  bijection onto the occupied voxels (checked through a render that reads
  every hit voxel's albedo by its index).
"""
import numpy as np
import pytest

import oracle_ffi as O
import vt
from test_gpu_parity import compare
from vt import scenes as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(ctx):
    print("[full-size] generating the 1024^3 world", flush=True)
    grid = S.generate(S.spec("c2"))
    pal = S.palette()
    world = vt.World.dense(ctx, grid, pal)
    yield grid, pal, world
    world.close()


def _render(ctx, world, cam, W, H, flags=0, traversal=vt.VT_TRAVERSAL_FLAT, stats=False,
            gbuffer=False, **kw):
    f = vt.Frame.alloc(W, H, gbuffer=gbuffer, steps=True, stats=stats)
    vt.render(ctx, world, cam, f, vt.params(W, H, S.MAX_STEPS["c2"], S.MAX_STEPS["c2"], S.SUN,
                                            flags, traversal, **kw))
    return f


def _dict(f):
    return {
        "shaded": f.shaded.cpu().numpy(),
        "hits": f.hits.cpu().numpy().view(np.uint32),
        "color": f.color.cpu().numpy(),
        "normal": f.normal.cpu().numpy(),
        "position": f.position.cpu().numpy(),
        "illumination": f.illumination.cpu().numpy(),
        "steps": f.steps.cpu().numpy().view(np.uint32),
    }


def test_c2_views_vs_oracle_bands(ctx, c2):
    import torch
    grid, pal, world = c2
    ow = O.DenseWorld(grid, pal)
    views = S.views("c2")
    for i in (0, 17, 40, 63):
        cam, W, H = views[i]
        g = _dict(_render(ctx, world, cam, W, H, stats=True, gbuffer=True))
        g["fast"] = _dict(_render(ctx, world, cam, W, H, gbuffer=True))
        torch.cuda.synchronize()
        y0 = (H // 2) & ~7
        rows = (y0, min(H, y0 + 8))
        o = O.render(ow, cam, W, H, S.MAX_STEPS["c2"], S.MAX_STEPS["c2"], S.SUN, rows=rows)
        compare(g, o, rows=rows, stats=False, label=f"c2 view {i} {W}x{H} rows {rows}")


def test_c2_batch_and_traversals_equal_plain(ctx, c2):
    import torch
    grid, pal, world = c2
    views = S.views("c2")
    plain = []
    for cam, W, H in views:
        plain.append(_render(ctx, world, cam, W, H))
    frames = [vt.Frame.alloc(W, H, steps=True) for _, W, H in views]
    prm = vt.params(0, 0, S.MAX_STEPS["c2"], S.MAX_STEPS["c2"], S.SUN)
    print("[full-size] c2 batch of 64 views", flush=True)
    b = vt.Batch(ctx, world, [(cam, f) for (cam, _, _), f in zip(views, frames)], prm,
                 vt.tuning(chunk=2))
    b.launch()
    torch.cuda.synchronize()
    for i, (f, r) in enumerate(zip(frames, plain)):
        for k in ("hits", "steps", "shaded"):
            a = getattr(f, k).cpu().numpy().view(np.uint32)
            c = getattr(r, k).cpu().numpy().view(np.uint32)
            assert np.array_equal(a, c), f"c2 view {i}: batch {k} differs from vt_render"
    b.close()
    print("[full-size] c2 flat vs bitmask on every view", flush=True)
    for i, ((cam, W, H), r) in enumerate(zip(views, plain)):
        fb = _render(ctx, world, cam, W, H, traversal=vt.VT_TRAVERSAL_BITMASK)
        torch.cuda.synchronize()
        for k in ("hits", "steps"):
            a = getattr(fb, k).cpu().numpy().view(np.uint32)
            c = getattr(r, k).cpu().numpy().view(np.uint32)
            assert np.array_equal(a, c), f"c2 view {i}: bitmask {k} differs from flat"


def test_c4_lambert_batch_compact_albedo(ctx, c2):
    import torch
    grid, pal, world = c2
    n_occ = vt.build_compact_index(ctx, world)
    assert n_occ == int(np.count_nonzero(grid))
    # albedo of compact index i = (i >> 12, i & 4095, 1), both parts exact in
    # float32: a render that looks every hit voxel up by its index must find
    # its own rank in Morton order
    idx = torch.arange(n_occ, device="cuda", dtype=torch.int64)
    albedo = torch.stack([(idx >> 12).float(), (idx & 4095).float(),
                          torch.ones(n_occ, device="cuda")], 1).contiguous()
    flags = vt.VT_FLAG_LAMBERT | vt.VT_FLAG_ALBEDO_COMPACT
    views = S.views("c4")
    prm = vt.params(0, 0, S.MAX_STEPS["c4"], S.MAX_STEPS["c4"], S.SUN, flags, 0, 0.25,
                    albedo.view(-1))
    frames = [vt.Frame.alloc(W, H, steps=True) for _, W, H in views]
    print("[full-size] c4 Lambert batch, compact albedo", flush=True)
    b = vt.Batch(ctx, world, [(cam, f) for (cam, _, _), f in zip(views, frames)], prm,
                 vt.tuning(chunk=2))
    b.launch()
    torch.cuda.synchronize()
    for i in (0, 23, 47):
        cam, W, H = views[i]
        r = vt.Frame.alloc(W, H, steps=True)
        vt.render(ctx, world, cam, r, vt.params(W, H, S.MAX_STEPS["c4"], S.MAX_STEPS["c4"],
                                                 S.SUN, flags, 0, 0.25, albedo.view(-1)))
        torch.cuda.synchronize()
        f = frames[i]
        for k in ("hits", "steps", "shaded"):
            a = getattr(f, k).cpu().numpy().view(np.uint32)
            c = getattr(r, k).cpu().numpy().view(np.uint32)
            assert np.array_equal(a, c), f"c4 view {i}: batch {k} differs from vt_render"
        # shaded = albedo[rank] * factor with factor in [0.25, 1.25]: the blue
        # channel (albedo 1) is the factor, red / blue and green / blue give
        # the two parts of the rank
        hits = r.hits.cpu().numpy().view(np.uint32)
        s = r.shaded.cpu().numpy()
        hit = hits[..., 0] != 0xFFFFFFFF
        fac = s[..., 2][hit].astype(np.float64)
        assert np.all(fac > 0)
        hi = np.rint(s[..., 0][hit] / fac).astype(np.int64)
        lo = np.rint(s[..., 1][hit] / fac).astype(np.int64)
        assert lo.min() >= 0 and lo.max() < 4096
        rank = hi * 4096 + lo
        vox = hits[..., 0][hit].astype(np.int64)
        x, y, z = vox % 1024, (vox // 1024) % 1024, vox // (1024 * 1024)
        # the rank is the number of occupied voxels before this one in Morton
        # order: monotone in the Morton code, and 0 <= rank < n_occ
        assert rank.min() >= 0 and rank.max() < n_occ
        code = np.zeros_like(vox)
        for bit in range(10):
            code |= ((x >> bit) & 1) << (3 * bit)
            code |= ((y >> bit) & 1) << (3 * bit + 1)
            code |= ((z >> bit) & 1) << (3 * bit + 2)
        ucode, first = np.unique(code, return_index=True)
        assert np.all(np.diff(rank[first]) > 0), f"c4 view {i}: compact index not increasing"
    b.close()
