"""HIP render path vs the CPU restatement (oracle/), through the C-ABI.

Bar: hit records (voxel id, t bits, face/shadow/sub-voxel/material word),
step counts, G-buffer position/normal/illumination and the statistics
counters are bit-exact; shaded RGBA and sky colours within 1e-5 max-abs
(the only non-bit-exact operation is powf inside SkyDome2).
"""
import sys

import numpy as np
import pytest

import oracle_ffi as O
import vt
from vt import scenes as S

pytestmark = pytest.mark.gpu

RGB_TOL = 1e-5


def _frame_dict(f):
    out = {
        "shaded": f.shaded.cpu().numpy(),
        "hits": f.hits.cpu().numpy().view(np.uint32),
        "color": f.color.cpu().numpy(),
        "normal": f.normal.cpu().numpy(),
        "position": f.position.cpu().numpy(),
        "illumination": f.illumination.cpu().numpy(),
    }
    if f.steps is not None:
        out["steps"] = f.steps.cpu().numpy().view(np.uint32)
    if f.stats is not None:
        out["stats"] = f.stats.cpu().numpy().view(np.uint64)
    return out


POISON = 0x7FA5A5A5  # a NaN as float32; no hit record, step count or colour has it


def poison(f):
    """Fills every image of a frame with POISON before a render, so that a
    pixel the kernel leaves unwritten cannot pass on what the memory held
    (a frame's memory is often the previous frame's, of the same view)."""
    import torch
    for name in ("shaded", "hits", "color", "normal", "position", "illumination", "steps",
                 "display"):
        t = getattr(f, name, None)
        if t is not None:
            t.view(-1).view(torch.int32).fill_(POISON)
    return f


def gpu_render(ctx, world, cam, W, H, max_p, max_s, flags=0, traversal=0, ambient=0.0,
               albedo=None, sun=S.SUN):
    """Renders three times: with the statistics counters (the kernel variant
    that counts every lookup), without them (the variant the benchmark times,
    with its own traversal loops; storing step counts keeps the free-region
    test off), and without step counts with the free-region test forced on
    (rays end early where they can hit nothing more: every output but the
    step counts is unchanged).  compare() checks all three against the
    oracle."""
    import torch
    prm = vt.params(W, H, max_p, max_s, sun, flags, traversal, ambient, albedo)
    f = poison(vt.Frame.alloc(W, H, gbuffer=True, steps=True, stats=True))
    vt.render(ctx, world, cam, f, prm)
    ff = poison(vt.Frame.alloc(W, H, gbuffer=True, steps=True))
    vt.render(ctx, world, cam, ff, prm)
    prm_free = vt.params(W, H, max_p, max_s, sun, flags | vt.VT_FLAG_FREE_MAP_ON, traversal,
                         ambient, albedo)
    fr = poison(vt.Frame.alloc(W, H, gbuffer=True))
    vt.render(ctx, world, cam, fr, prm_free)
    torch.cuda.synchronize()
    out = _frame_dict(f)
    out["fast"] = _frame_dict(ff)
    out["free"] = _frame_dict(fr)
    return out


def compare(g, o, rows=None, stats=True, label=""):
    if "fast" in g:
        compare(g["fast"], o, rows=rows, stats=False, label=label + " [no-stats kernel]")
    if "free" in g:
        compare(g["free"], o, rows=rows, stats=False, label=label + " [free-region test]")
    sl = slice(None) if rows is None else slice(*rows)
    print(f"[parity] {label}", flush=True)
    gh, oh = g["hits"][sl], o["hits"][sl]
    bad = np.any(gh != oh, axis=-1)
    if bad.any():
        first = [(tuple(int(i) for i in ix), [hex(int(v)) for v in gh[tuple(ix)]],
                  [hex(int(v)) for v in oh[tuple(ix)]]) for ix in np.argwhere(bad)[:6]]
        print(f"[parity] {label}: (pixel, gpu record, oracle record) {first}", flush=True)
    assert not bad.any(), f"{label}: {bad.sum()} hit records differ, first at {np.argwhere(bad)[:5]}"
    if "steps" in g:
        assert np.array_equal(g["steps"][sl], o["steps"][sl]), f"{label}: step counts differ"
    for k in ("position", "normal", "illumination"):
        assert np.array_equal(g[k][sl].view(np.uint32), o[k][sl].view(np.uint32)), f"{label}: {k}"
    assert np.max(np.abs(g["color"][sl] - o["color"][sl])) <= RGB_TOL, f"{label}: color"
    err = np.max(np.abs(g["shaded"][sl] - o["shaded"][sl]))
    assert err <= RGB_TOL, f"{label}: shaded max-abs {err}"
    if stats:
        # steps, primary hits (= shadow rays), occluded, material reads, entity
        # hits; counters 4-5 (bitmask word fetches) are the kernel's own
        # cache bookkeeping and have no reference counterpart
        sel = [0, 1, 2, 3, 6, 7]
        assert np.array_equal(g["stats"][sel], o["stats"][sel]), \
            f"{label}: stats {g['stats']} vs {o['stats']}"


@pytest.fixture(scope="module")
def c0():
    grid = S.generate(S.spec("c0"))
    return grid, S.palette()


@pytest.mark.parametrize("traversal", [vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8])
@pytest.mark.parametrize("flags", [0, vt.VT_FLAG_GBUF_UNORM8])
def test_c0_full_frame(ctx, c0, traversal, flags):
    grid, pal = c0
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    cam, W, H = S.views("c0")[0]
    g = gpu_render(ctx, w, cam, W, H, 512, 512, flags=flags, traversal=traversal)
    o = O.render(ow, cam, W, H, 512, 512, S.SUN, flags=flags, traversal=traversal)
    compare(g, o, label=f"c0 traversal={traversal} flags={flags}")
    w.close()


@pytest.mark.parametrize("sun", [(-0.5, 0.7, -0.2), (0.0, 1.0, 0.0), (0.6, 0.5, -0.6),
                                 (-0.3, -0.9, 0.1)])
def test_c0_sun_directions(ctx, c0, sun):
    """Shadow rays along other suns: components of either sign and zero (the
    0.001 substitution); the shadow pass has a separate loop for suns with
    no negative component."""
    grid, pal = c0
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    cam, W, H = S.views("c0")[0]
    g = gpu_render(ctx, w, cam, W, H, 512, 512, sun=sun)
    o = O.render(ow, cam, W, H, 512, 512, sun)
    compare(g, o, label=f"c0 sun={sun}")
    w.close()
    # the same world in the engine's layout (its shadow rays take the
    # reference-layout kernel's loops)
    import brickmap_util as BU
    dim, chunks, pool, atlas = BU.embedding_a(grid, pal)
    wb = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    gb = gpu_render(ctx, wb, cam, W, H, 512, 512, sun=sun)
    compare(gb, o, stats=False, label=f"c0 reference layout sun={sun}")
    wb.close()


def test_c0_reference_step_caps(ctx, c0):
    """The reference's own caps: 192 primary, 48 shadow."""
    grid, pal = c0
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    cam, W, H = S.views("c0")[0]
    g = gpu_render(ctx, w, cam, W, H, 192, 48)
    o = O.render(ow, cam, W, H, 192, 48, S.SUN)
    compare(g, o, label="c0 caps 192/48")


def test_c1_rows_and_flat_equals_bitmask(ctx):
    grid = S.generate(S.spec("c1"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    cam, W, H = S.views("c1")[0]
    gf = gpu_render(ctx, w, cam, W, H, 2048, 2048, traversal=vt.VT_TRAVERSAL_FLAT)
    gb = gpu_render(ctx, w, cam, W, H, 2048, 2048, traversal=vt.VT_TRAVERSAL_BITMASK)
    # full frame: flat and bitmask traversals return identical hits
    for k in ("hits", "steps", "position", "normal", "illumination"):
        assert np.array_equal(gf[k].view(np.uint32), gb[k].view(np.uint32)), k
        assert np.array_equal(gf["fast"][k].view(np.uint32), gb["fast"][k].view(np.uint32)), k
    assert np.array_equal(gf["stats"][:4], gb["stats"][:4])
    # oracle on a band of rows through the horizon and the near terrain
    ow = O.DenseWorld(grid, pal)
    for rows in ((0, 24), (520, 560), (1056, 1080)):
        o = O.render(ow, cam, W, H, 2048, 2048, S.SUN, rows=rows, gbuffer=True)
        compare(gf, o, rows=rows, stats=False, label=f"c1 rows {rows}")
    w.close()


# ---------------------------------------------------------------------------
# reference layout: chunk table + block pool + model atlas (drop-in path)
# ---------------------------------------------------------------------------

def test_brickmap_embedding_a_matches_oracle_and_dense(ctx, c0):
    import brickmap_util as BU
    grid, pal = c0
    dim, chunks, pool, atlas = BU.embedding_a(grid, pal)
    w = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    cam, W, H = S.views("c0")[0]
    g = gpu_render(ctx, w, cam, W, H, 512, 512)
    o = O.render(O.BrickmapWorld(dim, chunks, pool, atlas), cam, W, H, 512, 512, S.SUN)
    compare(g, o, stats=False, label="brickmap embedding A")
    wd = vt.World.dense(ctx, grid, pal)
    gd = gpu_render(ctx, wd, cam, W, H, 512, 512)
    for k in ("hits", "steps", "position", "normal", "illumination"):
        assert np.array_equal(g[k].view(np.uint32), gd[k].view(np.uint32)), k
        assert np.array_equal(g["fast"][k].view(np.uint32), gd["fast"][k].view(np.uint32)), k
    w.close()
    wd.close()


@pytest.mark.parametrize("flags", [0, vt.VT_FLAG_GBUF_UNORM8])
def test_brickmap_embedding_c_two_levels(ctx, c0, flags):
    import brickmap_util as BU
    grid, pal = c0
    dim, chunks, pool, atlas = BU.embedding_c(grid, pal)
    w = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    cam = vt.look_at((8.0, 13.75, -2.5), (8.0, 8.0, 8.0), 70.0)
    W = H = 256
    g = gpu_render(ctx, w, cam, W, H, 1024, 256, flags=flags)
    o = O.render(O.BrickmapWorld(dim, chunks, pool, atlas), cam, W, H, 1024, 256, S.SUN,
                 flags=flags)
    compare(g, o, stats=False, label=f"brickmap embedding C flags={flags}")
    assert (g["hits"][..., 0] != 0xFFFFFFFF).mean() > 0.3
    w.close()


# ---------------------------------------------------------------------------
# present pass (blit.fragment.glsl:32-35): crosshair + vignette + rgba8 store
# ---------------------------------------------------------------------------

def _display_close(g, o, label):
    """RGBA8 bytes equal, except +-1 where the GPU's powf and the host's
    differ across a rounding boundary (counted, at most 1e-4 of pixels)."""
    gb = g.view(np.uint8).astype(np.int16)
    ob = o.view(np.uint8).astype(np.int16)
    diff = np.abs(gb - ob)
    assert diff.max() <= 1, f"{label}: display bytes differ by {diff.max()}"
    assert np.count_nonzero(diff) <= max(1, diff.size // 10000), f"{label}: {np.count_nonzero(diff)}"


@pytest.mark.parametrize("flags", [vt.VT_FLAG_PRESENT, vt.VT_FLAG_PRESENT | vt.VT_FLAG_GBUF_UNORM8])
def test_c0_present_fused(ctx, c0, flags):
    import torch
    grid, pal = c0
    w = vt.World.dense(ctx, grid, pal)
    cam, W, H = S.views("c0")[0]
    f = vt.Frame.alloc(W, H, display=True)
    print(f"[parity] c0 present flags={flags}", flush=True)
    vt.render(ctx, w, cam, f, vt.params(W, H, 512, 512, S.SUN, flags))
    torch.cuda.synchronize()
    o = O.render(O.DenseWorld(grid, pal), cam, W, H, 512, 512, S.SUN, flags=flags, gbuffer=False,
                 display=True)
    assert np.array_equal(f.hits.cpu().numpy().view(np.uint32), o["hits"])
    err = np.max(np.abs(f.shaded.cpu().numpy() - o["shaded"]))
    assert err <= RGB_TOL, f"presented colour max-abs {err}"
    _display_close(f.display.cpu().numpy().view(np.uint32), o["display"], "c0 display")
    # the present pass of the plain composite equals the fused result
    f2 = vt.Frame.alloc(W, H)
    vt.render(ctx, w, cam, f2, vt.params(W, H, 512, 512, S.SUN, flags & ~vt.VT_FLAG_PRESENT))
    out = torch.empty_like(f2.shaded)
    disp = torch.empty((H, W), dtype=torch.int32, device="cuda")
    vt.present(ctx, f2.shaded, out, disp)
    torch.cuda.synchronize()
    assert torch.equal(out, f.shaded) and torch.equal(disp, f.display)
    w.close()


def test_present_standalone_1080p(ctx):
    import torch
    rng = np.random.default_rng(5)
    comp = rng.uniform(-0.2, 1.2, (1080, 1920, 4)).astype(np.float32)
    c = torch.from_numpy(comp).cuda()
    out = torch.empty_like(c)
    disp = torch.empty((1080, 1920), dtype=torch.int32, device="cuda")
    print("[parity] vt_present 1920x1080", flush=True)
    vt.present(ctx, c, out, disp)
    torch.cuda.synchronize()
    oo, od = O.present(comp)
    assert np.max(np.abs(out.cpu().numpy() - oo)) <= RGB_TOL
    _display_close(disp.cpu().numpy().view(np.uint32), od, "1080p display")
    # display only
    d2 = torch.zeros_like(disp)
    vt.present(ctx, c, None, d2)
    torch.cuda.synchronize()
    assert torch.equal(d2, disp)


def test_present_argument_errors(ctx):
    import torch
    c = torch.zeros((4, 4, 4), device="cuda")
    with pytest.raises(vt.VtError) as e:
        vt.present(ctx, c, None, None)
    assert e.value.status == vt._native.VT_ERR_NULL_POINTER
    grid = S.generate(S.spec("c0"))
    w = vt.World.dense(ctx, grid, S.palette())
    cam, W, H = S.views("c0")[0]
    f = vt.Frame.alloc(W, H, display=True)
    with pytest.raises(vt.VtError) as e:  # display without VT_FLAG_PRESENT
        vt.render(ctx, w, cam, f, vt.params(W, H, 512, 512, S.SUN, 0))
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    with pytest.raises(vt.VtError) as e:  # PRESENT with LAMBERT
        alb = torch.zeros(128 ** 3 * 3, device="cuda")
        vt.render(ctx, w, cam, f, vt.params(W, H, 512, 512, S.SUN,
                                             vt.VT_FLAG_PRESENT | vt.VT_FLAG_LAMBERT, 0, 0.2, alb))
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    w.close()
