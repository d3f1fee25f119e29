"""Batched persistent rendering (vt_batch_*) vs one vt_render per view.

The batch runs the same per-pixel path through a work queue, so every view's
outputs must be bit-identical to a plain render of that view, for every tile
shape, queue chunk and traversal; a captured hipGraph must reproduce the
eager launch on every replay (with other work and synchronisations between
replays)."""
import numpy as np
import pytest

import vt
from vt import scenes as S

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (123, 77), (8, 8), (3, 5), (300, 17), (64, 200)]


def make_views():
    rng = np.random.default_rng(11)
    out = [S.views("c0")[0][0]]
    for _ in range(len(SIZES) - 1):
        eye = rng.uniform(-30, 160, 3)
        eye[1] = rng.uniform(90, 140)
        out.append(vt.look_at(eye, rng.uniform(20, 108, 3), 70.0))
    return out


@pytest.fixture(scope="module")
def setup(ctx):
    import torch
    grid = S.generate(S.spec("c0"))
    world = vt.World.dense(ctx, grid, S.palette())
    cams = make_views()
    prm = vt.params(0, 0, 512, 512)
    ref = []
    for cam, (W, H) in zip(cams, SIZES):
        f = vt.Frame.alloc(W, H, steps=True, stats=True)
        vt.render(ctx, world, cam, f, vt.params(W, H, 512, 512))
        ref.append(f)
    torch.cuda.synchronize()
    yield world, cams, prm, ref
    world.close()


def fresh_frames():
    fs = [vt.Frame.alloc(W, H, steps=True) for (W, H) in SIZES]
    for f in fs:
        f.shaded.fill_(float("nan"))
        f.hits.fill_(-7)
        f.steps.fill_(-7)
    return fs


def assert_same(frames, ref, label):
    for i, (f, r) in enumerate(zip(frames, ref)):
        for k in ("hits", "steps", "shaded"):
            a = getattr(f, k).cpu().numpy().view(np.uint32)
            b = getattr(r, k).cpu().numpy().view(np.uint32)
            assert np.array_equal(a, b), f"{label}: view {i} {SIZES[i]} {k} differs"


@pytest.mark.parametrize("tile", [vt.VT_TILE_8X8, vt.VT_TILE_16X4, vt.VT_TILE_32X2])
@pytest.mark.parametrize("chunk", [1, 4])
@pytest.mark.parametrize("traversal", [vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8])
def test_batch_matches_per_view_render(ctx, setup, tile, chunk, traversal):
    import torch
    world, cams, prm, ref = setup
    frames = fresh_frames()
    b = vt.Batch(ctx, world, list(zip(cams, frames)), prm, vt.tuning(tile, chunk, traversal))
    b.launch()
    torch.cuda.synchronize()
    print(f"[batch] tile={tile} chunk={chunk} traversal={traversal}", flush=True)
    assert_same(frames, ref, f"tile={tile} chunk={chunk} trav={traversal}")
    b.close()


def test_batch_stats_sum_views(ctx, setup):
    import torch
    world, cams, prm, ref = setup
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    frames = fresh_frames()
    b = vt.Batch(ctx, world, list(zip(cams, frames)), prm, vt.tuning(), stats=stats)
    b.launch()
    torch.cuda.synchronize()
    total = sum(r.stats.cpu().numpy() for r in ref)
    assert np.array_equal(stats.cpu().numpy()[[0, 1, 2, 3, 7]], total[[0, 1, 2, 3, 7]])
    b.close()


def test_batch_graph_replays(ctx, setup):
    import torch
    world, cams, prm, ref = setup
    frames = fresh_frames()
    b = vt.Batch(ctx, world, list(zip(cams, frames)), prm, vt.tuning(vt.VT_TILE_16X4, 2))
    b.capture_graph()
    side = vt.Frame.alloc(64, 64)
    for rep in range(4):
        for f in frames:  # clear the outputs with the program's own eager work
            f.shaded.fill_(float("nan"))
            f.hits.fill_(-7)
            f.steps.fill_(-7)
        vt.render(ctx, world, cams[1], side, vt.params(64, 64, 512, 512))
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        print(f"[batch] graph replay {rep}", flush=True)
        assert_same(frames, ref, f"graph replay {rep}")
    b.close()


def test_autotune_picks_a_valid_tuning(ctx, setup):
    world, cams, prm, ref = setup
    frames = fresh_frames()
    best, ms = vt.autotune(ctx, world, list(zip(cams, frames)), prm, reps=1)
    assert best.tile in (3, 4, 5) and best.chunk in (1, 2, 4, 8)
    assert best.traversal in (0, 1, 2) and best.blocks_per_cu == 8
    assert 0.0 < ms < 1000.0
    # the frames hold a render when it returns
    assert_same(frames, ref, "after autotune")


def test_autotune_rejects_gradient_accumulation(ctx, setup):
    """VT_FLAG_GRAD_L2 would add into the caller's accumulators on every
    timed launch: refused before any launch (the accumulators stay zero)."""
    import torch
    world, cams, _, _ = setup
    frames = fresh_frames()
    for f in frames:
        f.target = torch.zeros_like(f.shaded)
    alb = torch.full((128 ** 3 * 3,), 0.5, device="cuda")
    ga = torch.zeros_like(alb)
    gs = torch.zeros(3, device="cuda")
    loss = torch.zeros(1, device="cuda")
    prm = vt.params(0, 0, 512, 512, S.SUN, vt.VT_FLAG_LAMBERT | vt.VT_FLAG_GRAD_L2, 0, 0.2, alb,
                    ga, gs, loss)
    with pytest.raises(vt.VtError) as e:
        vt.autotune(ctx, world, list(zip(cams, frames)), prm, reps=1)
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    torch.cuda.synchronize()
    assert not ga.any() and not gs.any() and not loss.any()


def test_batch_present_matches_per_view(ctx, setup):
    """VT_FLAG_PRESENT in a batch: each view presented with its own size."""
    import torch
    world, cams, _, _ = setup
    flags = vt.VT_FLAG_PRESENT | vt.VT_FLAG_GBUF_UNORM8
    ref, frames = [], []
    for cam, (W, H) in zip(cams, SIZES):
        f = vt.Frame.alloc(W, H, display=True)
        vt.render(ctx, world, cam, f, vt.params(W, H, 512, 512, S.SUN, flags))
        ref.append(f)
        frames.append(vt.Frame.alloc(W, H, display=True))
    b = vt.Batch(ctx, world, list(zip(cams, frames)), vt.params(0, 0, 512, 512, S.SUN, flags),
                 vt.tuning(vt.VT_TILE_16X4, 2))
    b.launch()
    torch.cuda.synchronize()
    print("[batch] present", flush=True)
    for i, (f, r) in enumerate(zip(frames, ref)):
        assert torch.equal(f.shaded, r.shaded) and torch.equal(f.display, r.display), i
    b.close()


@pytest.mark.parametrize("compact", [False, True])
def test_batch_fused_gradient_sums_views(ctx, setup, compact):
    """LAMBERT | GRAD_L2 in one batch launch: the loss and gradients are the
    sums of the per-view fused renders (float atomics: rtol 1e-5)."""
    import torch
    world, cams, _, _ = setup
    n = 128
    flags = vt.VT_FLAG_LAMBERT | vt.VT_FLAG_GRAD_L2
    nalb = n ** 3
    if compact:
        nalb = vt.build_compact_index(ctx, world)
        flags |= vt.VT_FLAG_ALBEDO_COMPACT
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    albedo = torch.rand(nalb * 3, generator=g, device="cuda")
    targets = [torch.rand((H, W, 4), generator=g, device="cuda") for (W, H) in SIZES]

    def grads():
        return (torch.zeros_like(albedo), torch.zeros(3, device="cuda"),
                torch.zeros(1, device="cuda"))

    ga, gs, gl = grads()
    for cam, (W, H), t in zip(cams, SIZES, targets):
        f = vt.Frame.alloc(W, H, hits=False)
        f.target = t
        vt.render(ctx, world, cam, f, vt.params(W, H, 512, 512, S.SUN, flags, 0, 0.2, albedo,
                                                ga, gs, gl))
    ba, bs, bl = grads()
    frames = []
    for (W, H), t in zip(SIZES, targets):
        f = vt.Frame.alloc(W, H, hits=False)
        f.target = t
        frames.append(f)
    b = vt.Batch(ctx, world, list(zip(cams, frames)),
                 vt.params(0, 0, 512, 512, S.SUN, flags, 0, 0.2, albedo, ba, bs, bl),
                 vt.tuning())
    b.launch()
    torch.cuda.synchronize()
    print(f"[batch] fused gradient compact={compact}", flush=True)
    assert torch.allclose(ba, ga, rtol=1e-5, atol=1e-5)
    assert torch.allclose(bl, gl, rtol=1e-5)
    assert torch.allclose(bs, gs, rtol=1e-4, atol=1e-2)
    assert float(gl) > 0 and torch.count_nonzero(ga) > 1000
    b.close()
