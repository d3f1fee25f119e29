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
from test_gpu_parity import POISON, compare, poison
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
            gbuffer=False, steps=True, **kw):
    f = poison(vt.Frame.alloc(W, H, gbuffer=gbuffer, steps=steps, stats=stats))
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
        **({} if f.steps is None else {"steps": f.steps.cpu().numpy().view(np.uint32)}),
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
        # no step counts: the free-region test ends rays early
        g["free"] = _dict(_render(ctx, world, cam, W, H, gbuffer=True, steps=False))
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
    frames = [poison(vt.Frame.alloc(W, H, steps=True)) for _, W, H in views]
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
    # frames without step counts: the batch (and vt_render) end rays early in
    # free regions; every other output is unchanged
    print("[full-size] c2 batch of 64 views, free-region test", flush=True)
    frames = [poison(vt.Frame.alloc(W, H)) for _, W, H in views]
    b = vt.Batch(ctx, world, [(cam, f) for (cam, _, _), f in zip(views, frames)], prm,
                 vt.tuning(chunk=2))
    b.launch()
    torch.cuda.synchronize()
    for i, ((cam, W, H), f, r) in enumerate(zip(views, frames, plain)):
        rf = _render(ctx, world, cam, W, H, steps=False)
        torch.cuda.synchronize()
        for k in ("hits", "shaded"):
            c = getattr(r, k).cpu().numpy().view(np.uint32)
            for src, g in (("batch", f), ("vt_render", rf)):
                a = getattr(g, k).cpu().numpy().view(np.uint32)
                assert np.array_equal(a, c), f"c2 view {i}: {src} {k} with the free-region test"
    b.close()
    print("[full-size] c2 flat vs bitmask (4^3 and 8^3 bricks) on every view", flush=True)
    for i, ((cam, W, H), r) in enumerate(zip(views, plain)):
        for trav in (vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8):
            fb = _render(ctx, world, cam, W, H, traversal=trav)
            torch.cuda.synchronize()
            for k in ("hits", "steps"):
                a = getattr(fb, k).cpu().numpy().view(np.uint32)
                c = getattr(r, k).cpu().numpy().view(np.uint32)
                assert np.array_equal(a, c), f"c2 view {i}: traversal {trav} {k} differs from flat"


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
    frames = [poison(vt.Frame.alloc(W, H, steps=True)) for _, W, H in views]
    print("[full-size] c4 Lambert batch, compact albedo", flush=True)
    b = vt.Batch(ctx, world, [(cam, f) for (cam, _, _), f in zip(views, frames)], prm,
                 vt.tuning(chunk=2))
    b.launch()
    torch.cuda.synchronize()
    # the same batch without step counts (free-region test on): same images
    frames_free = [poison(vt.Frame.alloc(W, H)) for _, W, H in views]
    bf = vt.Batch(ctx, world, [(cam, f) for (cam, _, _), f in zip(views, frames_free)], prm,
                  vt.tuning(chunk=2))
    bf.launch()
    torch.cuda.synchronize()
    for i, (f, ff) in enumerate(zip(frames, frames_free)):
        for k in ("hits", "shaded"):
            a = getattr(ff, k).cpu().numpy().view(np.uint32)
            c = getattr(f, k).cpu().numpy().view(np.uint32)
            assert np.array_equal(a, c), f"c4 view {i}: {k} with the free-region test"
    bf.close()
    del frames_free
    for i in (0, 23, 47):
        cam, W, H = views[i]
        r = poison(vt.Frame.alloc(W, H, steps=True))
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


def test_c2_repeated_renders_identical(ctx, c2):
    """The same large views rendered again and again, into frames filled
    with POISON first: every render writes every pixel and gives the same
    records, with step counts stored (free-region test off) and without
    (test on).  A kernel whose lanes exchange work through LDS once lost or
    mixed up some pixels' records from one render to the next here."""
    import torch
    grid, pal, world = c2
    views = S.views("c2")
    for i in (2, 35):
        cam, W, H = views[i]
        for steps in (True, False):
            ref = None
            for k in range(6):
                f = _render(ctx, world, cam, W, H, steps=steps)
                torch.cuda.synchronize()
                h = f.hits.cpu().numpy().view(np.uint32)
                assert not np.any(h == POISON), f"view {i} render {k}: pixels left unwritten"
                if ref is None:
                    ref = h.copy()
                else:
                    n = int(np.sum(np.any(h != ref, -1)))
                    assert n == 0, f"view {i} steps={steps} render {k}: {n} records changed"


# ---------------------------------------------------------------------------
# whole views against the oracle
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("part", range(4))
def test_c2_whole_views_vs_oracle(ctx, c2, part):
    """Every C2 view whole (16 views per part): the timed kernel's hit
    records, step counts and G-buffer bit-exact; the statistics kernel too
    on the views of at most 512x512 pixels."""
    import torch
    grid, pal, world = c2
    ow = O.DenseWorld(grid, pal)
    views = S.views("c2")
    cap = S.MAX_STEPS["c2"]
    n6 = n4 = 0
    for i in range(16 * part, 16 * part + 16):
        cam, W, H = views[i]
        g = {"fast": _dict(_render(ctx, world, cam, W, H, gbuffer=True))}
        small = W * H <= 512 * 512
        if small:
            g.update(_dict(_render(ctx, world, cam, W, H, stats=True, gbuffer=True)))
        torch.cuda.synchronize()
        o = O.render(ow, cam, W, H, cap, cap, S.SUN)
        m = o["tie_margin"]
        n6 += int((m < 1e-6).sum())
        n4 += int((m < 1e-4).sum())
        label = f"c2 view {i} {W}x{H} whole"
        if small:
            compare(g, o, stats=False, label=label)
        else:
            compare(g["fast"], o, stats=False, label=label + " [no-stats kernel]")
    print(f"[full-size] c2 views {16 * part}..{16 * part + 15}: tie margin < 1e-6 on {n6} "
          f"pixels, < 1e-4 on {n4} (compared bit for bit, none exempted)", flush=True)


# ---------------------------------------------------------------------------
# C4: the differentiable batch at full size
# ---------------------------------------------------------------------------

AMBIENT_C4 = 0.2


def test_c4_backward_full_size(ctx, c2):
    """48 views (20 M pixels), compact albedo (72.7 M occupied voxels),
    ground truth = material colour + N(0, 0.05), start 0.5: the fused
    recompute batch (VT_FLAG_GRAD_L2) against the stored-record backward
    (Lambert batch writing hit records, then vt_lambert_backward over all
    records), and both against the oracle's FP64 gradient: the loss and the
    sun gradient over every pixel, the albedo gradient of 64 random visible
    voxels."""
    import torch
    grid, pal, world = c2
    n_occ = vt.build_compact_index(ctx, world)
    flags = vt.VT_FLAG_LAMBERT | vt.VT_FLAG_ALBEDO_COMPACT
    views = S.views("c4")
    cap = S.MAX_STEPS["c4"]
    npix = sum(W * H for _, W, H in views)
    print(f"[full-size] c4: {len(views)} views, {npix} pixels, {n_occ} occupied voxels",
          flush=True)
    gt = S.ground_truth_albedo(grid, pal, compact=True)
    assert gt.numel() == 3 * n_occ
    start = torch.full_like(gt, 0.5)
    hits = torch.empty((npix, 3), dtype=torch.int32, device="cuda")
    shaded = torch.empty((npix, 4), dtype=torch.float32, device="cuda")
    target = torch.empty((npix, 4), dtype=torch.float32, device="cuda")

    def frames(**bufs):
        out, off = [], 0
        for cam, W, H in views:
            f = vt.Frame(W, H)
            for k, t in bufs.items():
                setattr(f, k, t[off:off + W * H].view(H, W, -1))
            out.append((cam, f))
            off += W * H
        return out

    def prm(alb, fl, grads=(None, None, None)):
        return vt.params(0, 0, cap, cap, S.SUN, fl, 0, AMBIENT_C4, alb, *grads)

    tune = vt.tuning(chunk=2)
    print("[full-size] c4 targets: Lambert batch of the ground truth", flush=True)
    b = vt.Batch(ctx, world, frames(shaded=target), prm(gt, flags), tune)
    b.launch()
    torch.cuda.synchronize()
    b.close()
    del gt
    print("[full-size] c4 stored records + backward", flush=True)
    b = vt.Batch(ctx, world, frames(shaded=shaded, hits=hits), prm(start, flags), tune)
    b.launch()
    b.close()
    ga_s = torch.zeros_like(start)
    gs_s = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss_s = torch.zeros(1, dtype=torch.float32, device="cuda")
    vt.lambert_backward(ctx, world, hits, vt.params(1, 1, sun=S.SUN, flags=flags,
                                                    ambient=AMBIENT_C4, albedo=start),
                        target, vt.VT_UPSTREAM_L2_TARGET, ga_s, gs_s, loss_s)
    print("[full-size] c4 fused recompute batch", flush=True)
    ga_f = torch.zeros_like(start)
    gs_f = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss_f = torch.zeros(1, dtype=torch.float32, device="cuda")
    shaded_f = torch.empty_like(shaded)
    b = vt.Batch(ctx, world, frames(shaded=shaded_f, target=target),
                 prm(start, flags | vt.VT_FLAG_GRAD_L2, (ga_f, gs_f, loss_f)), tune)
    b.launch()
    b.close()
    # per-entry scale of the float32 rounding error: the sum over pixels of
    # 2 fct (|I| + |T|) -- the operands of each term's residual I - T
    # (upstream 2 (|I| + |T|) through the upstream-gradient backward)
    bound = torch.zeros_like(start)
    up = (2.0 * (shaded.abs() + target.abs())).contiguous()
    vt.lambert_backward(ctx, world, hits, vt.params(1, 1, sun=S.SUN, flags=flags,
                                                    ambient=AMBIENT_C4, albedo=start),
                        up, vt.VT_UPSTREAM_GRADIENT, bound)
    torch.cuda.synchronize()
    assert torch.equal(shaded_f, shaded), "fused and stored forwards differ"
    diff = (ga_f - ga_s).abs()
    worst = float((diff - (2e-5 * bound + 2e-6)).max().item())
    print(f"[full-size] c4 loss stored {loss_s.item():.7g} fused {loss_f.item():.7g}; "
          f"grad worst excess {worst:.3g}", flush=True)
    assert worst <= 0.0
    assert abs(loss_f.item() - loss_s.item()) <= 2e-5 * abs(loss_s.item())

    # FP64 over every pixel: loss and sun gradient (the start albedo is 0.5
    # everywhere, so every hit record may name one albedo slot)
    rec = hits.cpu().numpy().view(np.uint32).copy()
    tgt = target.cpu().numpy()
    hit = rec[:, 0] != 0xFFFFFFFF
    rec1 = rec.copy()
    rec1[hit, 0] = 0
    half = np.full(3, 0.5, np.float32)
    da = np.zeros(3, np.float64)
    ds = np.zeros(3, np.float64)
    loss64 = O.lib().or_lambert_grad64(rec1.ctypes.data, npix, half.ctypes.data, O.fvec(S.SUN),
                                       AMBIENT_C4, tgt.ctypes.data, da.ctypes.data,
                                       ds.ctypes.data)
    l = np.asarray(S.SUN, np.float64)
    l /= np.linalg.norm(l)
    normals = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]],
                       np.float64)
    info = rec[hit, 2]
    nd = normals[(info & 7).astype(np.int64) - 1] @ l
    lit = (nd > 0) & (((info >> 3) & 1) == 0)
    fct = AMBIENT_C4 + np.maximum(nd, 0) * (((info >> 3) & 1) == 0)
    mag = 0.5 * fct[:, None] + np.abs(tgt[hit, :3].astype(np.float64))
    sb = float(np.sum(2 * mag * 0.5, 1)[lit].sum() * 2 / np.linalg.norm(S.SUN))
    for name, L_, gs in (("stored", loss_s, gs_s), ("fused", loss_f, gs_f)):
        g = gs.cpu().numpy()
        print(f"[full-size] c4 {name}: loss {L_.item():.7g} fp64 {loss64:.7g}; grad_sun {g} "
              f"fp64 {ds} (term bound {sb:.3g})", flush=True)
        assert abs(L_.item() - loss64) <= 2e-5 * abs(loss64)
        assert np.all(np.abs(g - ds) <= 1e-5 * sb), (name, g, ds, sb)

    # FP64 albedo gradient of 64 random visible voxels, through their compact
    # index (rank of the Morton code among the occupied voxels)
    rng = np.random.default_rng(44)
    vis = np.unique(rec[hit, 0])
    pick = rng.choice(vis, 64, replace=False).astype(np.int64)
    _, codes = S.occupied_morton(grid)
    pc = S.morton_codes(torch.from_numpy(pick).cuda(), grid.shape[0])
    rank = torch.searchsorted(codes, pc).cpu().numpy()
    assert np.array_equal(codes[torch.from_numpy(rank).cuda()].cpu().numpy(), pc.cpu().numpy())
    del codes
    slot = {int(v): j for j, v in enumerate(pick)}
    sel = np.isin(rec[:, 0], pick)
    sub = rec[sel].copy()
    sub[:, 0] = [slot[int(v)] for v in sub[:, 0]]
    sub_t = np.ascontiguousarray(tgt[sel])
    alb = np.full(64 * 3, 0.5, np.float32)
    da = np.zeros(64 * 3, np.float64)
    ds2 = np.zeros(3, np.float64)
    O.lib().or_lambert_grad64(sub.ctypes.data, sub.shape[0], alb.ctypes.data, O.fvec(S.SUN),
                              AMBIENT_C4, sub_t.ctypes.data, da.ctypes.data, ds2.ctypes.data)
    idx = torch.from_numpy((3 * rank[:, None] + np.arange(3)).reshape(-1)).cuda()
    for name, ga in (("stored", ga_s), ("fused", ga_f)):
        g = ga[idx].cpu().numpy().astype(np.float64)
        B = bound[idx].cpu().numpy().astype(np.float64)
        err = np.abs(g - da)
        assert np.all(err <= 1e-5 * B + 1e-6), (name, float(np.max(err - 1e-5 * B)))
    assert np.count_nonzero(da) >= 64
