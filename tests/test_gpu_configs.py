"""The benchmark and engine configurations at their own sizes, on the GPU,
against the oracle (whole frames, not bands).

* C1: the headline 1920x1080 frame of the 512^3 world, whole, for every
  traversal (flat; bitmask with 4^3 and with 8^3 bricks) and both kernel
  variants (the statistics
  kernel and the timed one): hit records, step counts and G-buffer images
  bit-exact, shaded within 1e-5, statistics counters equal.
* C3 at its configuration: 256^3 world (the C0 generator at 2x), 16 orbit
  views of 512x512; ground-truth albedo = material colour + N(0, 0.05),
  optimisation start 0.5.  Every view's hits against the oracle; the
  Lambert forward bit-exact against the oracle's shading of the same
  records; the stored-record backward over all 16 views against the
  oracle's FP64 gradient; central differences (h = 1e-3) on 64 random
  visible voxels and the 3 sun components; the fused recompute mode
  (VT_FLAG_GRAD_L2) against the stored-record backward.
* The engine's own frame configuration: its generated 512^3 world with the
  29 model slots of its .vox assets, Game.init's camera, the G-buffer size
  GBuffer.init(1280, 720) (game.zig:91), the shaders' step caps 192 / 48
  (primary.comp.glsl:43, secondary.comp.glsl:41), rgba8 G-buffer images
  (gbuffer.zig:10-13) and SUN_DIR (camera.glsl:8): whole frames.

Every comparison prints the number of pixels whose rays (primary or shadow)
met a crossing-time tie margin below 1e-6 and below 1e-4 (the oracle's
per-ray minimum of second-smallest minus smallest t over its DDA steps).
Ties are reproduced exactly (the z > y > x rule, map.glsl:158), so none of
these rays is exempted from the bit-exact bar.

Gradient tolerances (float32 arithmetic, atomics summing in a
nondeterministic order): per voxel and channel |g - g64| <= 1e-5 * B + 1e-6,
B = the FP64 sum over the voxel's pixels of 2 fct (|A fct| + |T|) -- the
magnitudes of the operands of each float32 term, since the residual
A fct - T is a rounded difference; the sun gradient within 1e-5 of the same
operand-magnitude sum of its terms; loss rtol 2e-5; central differences
1e-3 relative (albedo) and 1e-3 relative + 0.05 (sun).
"""
import math

import numpy as np
import pytest

import oracle_ffi as O
import vt
from test_gpu_parity import compare, gpu_render
from vt import _native as N
from vt import scenes as S

pytestmark = pytest.mark.gpu

AMBIENT = 0.2
SUN_DIR = (7.52185881e-01, 6.58950984e-01, 7.52185881e-01)  # camera.glsl:8


def log(msg):
    print(f"[configs] {msg}", flush=True)


def report_ties(label, o):
    m = o["tie_margin"]
    n6, n4 = int((m < 1e-6).sum()), int((m < 1e-4).sum())
    log(f"{label}: {m.size} pixels; tie margin < 1e-6: {n6}, < 1e-4: {n4} "
        f"(compared bit for bit, none exempted)")
    return n6, n4


# ---------------------------------------------------------------------------
# C1: the whole headline frame
# ---------------------------------------------------------------------------

def test_c1_whole_frame(ctx):
    grid = S.generate(S.spec("c1"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    cam, W, H = S.views("c1")[0]
    cap = S.MAX_STEPS["c1"]
    for trav, name in ((vt.VT_TRAVERSAL_FLAT, "flat"), (vt.VT_TRAVERSAL_BITMASK, "bitmask"),
                       (vt.VT_TRAVERSAL_BITMASK8, "bitmask8")):
        log(f"c1 oracle, whole frame ({name} statistics)")
        o = O.render(ow, cam, W, H, cap, cap, S.SUN, traversal=trav)
        if trav == vt.VT_TRAVERSAL_FLAT:
            report_ties("c1 1920x1080", o)
        g = gpu_render(ctx, w, cam, W, H, cap, cap, traversal=trav)
        compare(g, o, stats=True, label=f"c1 whole frame, {name}")
    w.close()


# ---------------------------------------------------------------------------
# C3 at its own configuration
# ---------------------------------------------------------------------------

def _fct64(info, sun):
    """ambient + max(n . l, 0) * vis per record, float64."""
    l = np.asarray(sun, np.float64)
    l = l / np.linalg.norm(l)
    face = (info & 7).astype(np.int64)
    normals = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]],
                       np.float64)
    nd = normals[np.clip(face - 1, 0, 5)] @ l
    vis = ((info >> 3) & 1) == 0
    return AMBIENT + np.maximum(nd, 0.0) * vis, nd, normals[np.clip(face - 1, 0, 5)]


@pytest.fixture(scope="module")
def c3(ctx):
    import torch
    log("c3: generating the 256^3 world")
    grid = S.generate(S.spec("c3"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    views = S.views("c3")
    assert len(views) == 16 and all((W, H) == (512, 512) for _, W, H in views)
    gt = S.ground_truth_albedo(grid, pal)
    gt_np = gt.cpu().numpy()
    start = torch.full_like(gt, 0.5)
    npix = sum(W * H for _, W, H in views)
    hits = torch.empty((npix, 3), dtype=torch.int32, device="cuda")
    target = torch.empty((npix, 4), dtype=torch.float32, device="cuda")
    cap = S.MAX_STEPS["c3"]
    off = 0
    frames = []
    for cam, W, H in views:
        f = vt.Frame(W, H)
        f.hits = hits[off:off + W * H].view(H, W, 3)
        f.shaded = target[off:off + W * H].view(H, W, 4)
        # targets: the Lambert render of the ground truth
        vt.render(ctx, w, cam, f, vt.params(W, H, cap, cap, S.SUN, vt.VT_FLAG_LAMBERT, 0,
                                            AMBIENT, gt))
        frames.append(f)
        off += W * H
    torch.cuda.synchronize()
    yield dict(grid=grid, pal=pal, world=w, views=views, gt_np=gt_np, gt=gt, start=start,
               hits=hits, target=target, npix=npix, cap=cap, frames=frames)
    w.close()


def test_c3_every_view_vs_oracle(ctx, c3):
    import torch
    ow = O.DenseWorld(c3["grid"], c3["pal"])
    cap = c3["cap"]
    hits_all = c3["hits"].cpu().numpy().view(np.uint32)
    tgt_all = c3["target"].cpu().numpy()
    off = 0
    n6 = n4 = 0
    for k, (cam, W, H) in enumerate(c3["views"]):
        o = O.render(ow, cam, W, H, cap, cap, S.SUN)
        a, b = report_ties(f"c3 view {k}", o)
        n6, n4 = n6 + a, n4 + b
        g = gpu_render(ctx, c3["world"], cam, W, H, cap, cap)
        compare(g, o, stats=True, label=f"c3 view {k}")
        # the Lambert targets were rendered with the same hit structure
        h = hits_all[off:off + W * H]
        assert np.array_equal(h, o["hits"].reshape(-1, 3)), f"c3 view {k}: Lambert-pass hits"
        ref = np.zeros((W * H, 4), np.float32)
        O.lib().or_shade_lambert(o["hits"].ctypes.data, W * H, c3["gt_np"].ctypes.data,
                                 O.fvec(S.SUN), AMBIENT, ref.ctypes.data)
        assert np.array_equal(tgt_all[off:off + W * H].view(np.uint32), ref.view(np.uint32)), \
            f"c3 view {k}: Lambert forward"
        off += W * H
    log(f"c3 all 16 views: tie margin < 1e-6 on {n6} pixels, < 1e-4 on {n4}")
    torch.cuda.synchronize()


def _backward(ctx, c3, albedo, sun=S.SUN):
    import torch
    ga = torch.zeros_like(albedo)
    gs = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    prm = vt.params(1, 1, sun=sun, flags=vt.VT_FLAG_LAMBERT, ambient=AMBIENT, albedo=albedo)
    vt.lambert_backward(ctx, c3["world"], c3["hits"], prm, c3["target"],
                        vt.VT_UPSTREAM_L2_TARGET, ga, gs, loss)
    torch.cuda.synchronize()
    return ga.cpu().numpy(), gs.cpu().numpy(), float(loss.item())


def _bounds64(c3, albedo_np, sun=S.SUN):
    """Scales of the float32 rounding error of the gradients, FP64: per
    albedo entry the sum over its pixels of 2 fct (|A fct| + |T|) -- the
    residual A fct - T is a float32 difference, so its error scales with its
    operands, not with the residual -- and for the sun the same operand
    magnitudes times |A| (the sum of |sun terms| with |r| so replaced)."""
    rec = c3["hits"].cpu().numpy().view(np.uint32)
    T = c3["target"].cpu().numpy().astype(np.float64)
    hit = rec[:, 0] != 0xFFFFFFFF
    v = rec[hit, 0].astype(np.int64)
    fct, nd, n = _fct64(rec[hit, 2], sun)
    A = albedo_np.reshape(-1, 3)[v].astype(np.float64)
    mag = np.abs(A * fct[:, None]) + np.abs(T[hit, :3])
    B = np.zeros(albedo_np.size, np.float64)
    for c in range(3):
        B[c::3] = np.bincount(v, 2 * mag[:, c] * fct, minlength=albedo_np.size // 3)
    lit = (nd > 0) & (((rec[hit, 2] >> 3) & 1) == 0)
    gsum = np.sum(2 * mag * np.abs(A), 1)
    sb = float(np.sum(gsum[lit]) * 2.0 / np.linalg.norm(np.asarray(sun, np.float64)))
    return B, sb


def test_c3_backward_vs_fp64(ctx, c3):
    rec = c3["hits"].cpu().numpy().view(np.uint32)
    tgt = c3["target"].cpu().numpy()
    start_np = c3["start"].cpu().numpy()
    log("c3 stored-record backward over 16 views vs FP64")
    ga, gs, loss = _backward(ctx, c3, c3["start"])
    da = np.zeros(start_np.size, np.float64)
    ds = np.zeros(3, np.float64)
    loss64 = O.lib().or_lambert_grad64(rec.ctypes.data, c3["npix"], start_np.ctypes.data,
                                       O.fvec(S.SUN), AMBIENT, tgt.ctypes.data, da.ctypes.data,
                                       ds.ctypes.data)
    B, sb = _bounds64(c3, start_np)
    log(f"loss {loss:.7g} fp64 {loss64:.7g}; grad_sun {gs} fp64 {ds} (term bound {sb:.3g})")
    assert abs(loss - loss64) <= 2e-5 * abs(loss64)
    assert np.count_nonzero(da) > 100000
    err = np.abs(ga.astype(np.float64) - da)
    assert np.all(err <= 1e-5 * B + 1e-6), f"worst {np.max(err - 1e-5 * B)}"
    assert np.all(np.abs(gs - ds) <= 1e-5 * sb), (gs, ds, sb)
    assert np.linalg.norm(ds) > 1e-4 * sb  # the tolerance is a small part of the gradient


def test_c3_finite_differences(ctx, c3):
    import torch
    ga, gs, _ = _backward(ctx, c3, c3["start"])
    T = c3["target"][:, :3].double()
    hits = c3["hits"]
    vid = hits[:, 0].cpu().numpy().view(np.uint32)
    visible = np.unique(vid[vid != 0xFFFFFFFF])
    rng = np.random.default_rng(11)
    picks = rng.choice(visible, 64, replace=False)

    def loss_of(albedo, sun=S.SUN):
        out = torch.empty((c3["npix"], 4), dtype=torch.float32, device="cuda")
        prm = vt.params(1, 1, sun=sun, flags=vt.VT_FLAG_LAMBERT, ambient=AMBIENT, albedo=albedo)
        vt.lambert_shade(ctx, c3["world"], hits, prm, out)
        return float(((out[:, :3].double() - T) ** 2).sum().item())

    h = 1e-3
    base = c3["start"]
    log("c3 central differences, 64 visible voxels")
    for v in picks:
        c = int(rng.integers(3))
        k = 3 * int(v) + c
        ap, am = base.clone(), base.clone()
        ap[k] += h
        am[k] -= h
        hp = float(ap[k].item()) - float(base[k].item())
        hm = float(base[k].item()) - float(am[k].item())
        fd = (loss_of(ap) - loss_of(am)) / (hp + hm)
        assert abs(fd - ga[k]) <= 1e-3 * max(1.0, abs(ga[k])), (int(v), c, fd, ga[k])
    log("c3 central differences, sun")
    sun = np.array(S.SUN, np.float32)
    for k in range(3):
        sp, sm = sun.copy(), sun.copy()
        sp[k] += np.float32(h)
        sm[k] -= np.float32(h)
        fd = (loss_of(base, sp) - loss_of(base, sm)) / float(sp[k] - sm[k])
        assert abs(fd - gs[k]) <= 1e-3 * abs(gs[k]) + 0.05, (k, fd, gs[k])


def test_c3_fused_recompute_matches_stored(ctx, c3):
    import torch
    ga2, gs2, loss2 = _backward(ctx, c3, c3["start"])
    B, sb = _bounds64(c3, c3["start"].cpu().numpy())
    ga = torch.zeros_like(c3["start"])
    gs = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    cap = c3["cap"]
    flags = vt.VT_FLAG_LAMBERT | vt.VT_FLAG_GRAD_L2
    log("c3 fused GRAD_L2 renders of the 16 views")
    for (cam, W, H), f in zip(c3["views"], c3["frames"]):
        r = vt.Frame.alloc(W, H)
        r.target = f.shaded
        vt.render(ctx, c3["world"], cam, r, vt.params(W, H, cap, cap, S.SUN, flags, 0, AMBIENT,
                                                     c3["start"], ga, gs, loss))
        torch.cuda.synchronize()
        assert torch.equal(r.hits, f.hits)
    torch.cuda.synchronize()
    err = np.abs(ga.cpu().numpy().astype(np.float64) - ga2)
    # two float32 sums of the same terms: twice the per-sum bound
    assert np.all(err <= 2e-5 * B + 2e-6)
    assert abs(loss.item() - loss2) <= 2e-5 * abs(loss2)
    assert np.all(np.abs(gs.cpu().numpy() - gs2) <= 2e-5 * sb)


# ---------------------------------------------------------------------------
# the engine's own frame configuration
# ---------------------------------------------------------------------------

def test_engine_frame_config(ctx):
    """Game.init's world and camera at the engine's G-buffer size, step caps,
    rgba8 G-buffer and SUN_DIR: whole frames, both kernel variants."""
    from test_gpu_edit import atlas_fixture
    from vt import camera as vcam
    atlas = atlas_fixture()
    rc, chunks, pool = O.procgen(512)
    assert rc == 0
    w = vt.World.empty_brickmap(ctx, 512, atlas)
    w.procgen()
    ow = O.BrickmapWorld(512, chunks, pool, atlas)
    start = N.Camera()
    start.position[:] = [256.0, 25.0, 256.0, 0.0]
    start.view[:] = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    start.fov = float(np.float32(math.pi / 2.0))
    turned = vcam.look_at((256.0, 25.0, 256.0), (180.0, 12.0, 330.0), 90.0)
    W, H = 1280, 720
    flags = vt.VT_FLAG_GBUF_UNORM8
    for k, cam in enumerate((start, turned)):
        log(f"engine frame {k}: 1280x720, caps 192/48, rgba8 G-buffer, SUN_DIR")
        o = O.render(ow, cam, W, H, 192, 48, SUN_DIR, flags=flags)
        report_ties(f"engine frame {k}", o)
        g = gpu_render(ctx, w, cam, W, H, 192, 48, flags=flags, sun=SUN_DIR)
        compare(g, o, stats=False, label=f"engine frame {k}")
        assert (g["hits"][..., 0] != 0xFFFFFFFF).mean() > 0.2
        # rgba8 images: the illumination alpha is 0 (occluded, -0.3 clamped)
        # or 0.3 rounded to 77/255
        a = np.unique(g["illumination"][..., 3])
        assert set(a.tolist()) <= {0.0, np.float32(77.0) / np.float32(255.0)}
    w.close()
