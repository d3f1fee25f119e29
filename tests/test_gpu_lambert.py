"""Differentiable Lambert x visibility path on the GPU, through the C-ABI.

Shading from hit records is project-defined (the reference has no
differentiable renderer): rgb = albedo[voxel] * (ambient + max(n.l, 0) * vis),
sky pixels 0.  Checks:

* forward: the fused render (VT_FLAG_LAMBERT), vt_lambert_shade over stored
  hit records and the oracle's or_shade_lambert agree bit for bit (same
  binary32 operations, no contraction);
* compact albedo: the occupancy prefix maps every occupied voxel to its rank
  in Morton order (every occupied voxel of the world is shaded once with the
  dense and once with the compacted table -- results bit-identical);
* backward: float32 atomics vs the oracle's FP64 gradient; central finite
  differences of the GPU forward (albedo, h = 1e-3; sun, h = 1e-3) against
  the GPU backward; the fused recompute mode (VT_FLAG_GRAD_L2) against the
  stored-record backward; the upstream-gradient kind through a
  torch.autograd.Function against the L2-target kind.

Tolerances (atomics make summation order nondeterministic): per-voxel albedo
gradients rtol 1e-5 + atol 1e-5 (terms are O(1)); loss rtol 2e-5; sun
gradient within 1e-5 x (a bound on the sum of |per-pixel terms|).
"""
import numpy as np
import pytest

import oracle_ffi as O
import vt
from vt import _native as N
from vt import scenes as S

pytestmark = pytest.mark.gpu

AMBIENT = 0.2


def log(msg):
    print(f"[lambert] {msg}", flush=True)


def part1by2(v):
    v = v.astype(np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0xFF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def compact_rank(grid):
    """(occupied voxel ids, rank of each in Morton order)."""
    dim = grid.shape[0]
    lg = int(dim).bit_length() - 1
    occ = np.flatnonzero(grid.ravel()).astype(np.uint64)
    m = np.uint64(dim - 1)
    x, y, z = occ & m, (occ >> np.uint64(lg)) & m, occ >> np.uint64(2 * lg)
    code = part1by2(x) | (part1by2(y) << np.uint64(1)) | (part1by2(z) << np.uint64(2))
    order = np.argsort(code, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return occ.astype(np.int64), rank


@pytest.fixture(scope="module")
def scene(ctx):
    import torch
    grid = S.generate(S.spec("c0"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    n = grid.shape[0]
    rng = np.random.default_rng(1234)
    albedo_np = rng.uniform(0.05, 1.0, n ** 3 * 3).astype(np.float32)
    albedo = torch.from_numpy(albedo_np).cuda()
    cam, W, H = S.views("c0")[0]
    f = vt.Frame.alloc(W, H, gbuffer=False)
    log("render c0 hit records")
    vt.render(ctx, w, cam, f, vt.params(W, H, 512, 512, S.SUN))
    torch.cuda.synchronize()
    hits = f.hits.contiguous()
    target = torch.from_numpy(rng.uniform(0, 1, (H, W, 4)).astype(np.float32)).cuda()
    yield dict(grid=grid, pal=pal, world=w, albedo_np=albedo_np, albedo=albedo, cam=cam, W=W,
               H=H, hits=hits, target=target)
    w.close()


def shade(ctx, sc, albedo, sun=S.SUN, flags=0, hits=None):
    import torch
    hits = sc["hits"] if hits is None else hits
    npix = hits.numel() // 3
    out = torch.empty((npix, 4), dtype=torch.float32, device="cuda")
    prm = vt.params(1, 1, sun=sun, flags=vt.VT_FLAG_LAMBERT | flags, ambient=AMBIENT,
                    albedo=albedo)
    vt.lambert_shade(ctx, sc["world"], hits, prm, out)
    torch.cuda.synchronize()
    return out


def backward(ctx, sc, albedo, upstream, kind, sun=S.SUN, flags=0, nalb=None):
    import torch
    nalb = albedo.numel() if nalb is None else nalb
    ga = torch.zeros(nalb, dtype=torch.float32, device="cuda")
    gs = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    prm = vt.params(1, 1, sun=sun, flags=vt.VT_FLAG_LAMBERT | flags, ambient=AMBIENT,
                    albedo=albedo)
    vt.lambert_backward(ctx, sc["world"], sc["hits"], prm, upstream, kind, ga, gs, loss)
    torch.cuda.synchronize()
    return ga.cpu().numpy(), gs.cpu().numpy(), float(loss.item())


def sun_term_bound(sc):
    """Bound on sum over lit pixels of |dL/dsun term|: |sum_c 2 r_c A_c| * 2 / |sun|."""
    hits = sc["hits"].cpu().numpy().view(np.uint32).reshape(-1, 3)
    tgt = sc["target"].cpu().numpy().reshape(-1, 4)
    I = oracle_shade(sc, sc["albedo_np"])
    hit = hits[:, 0] != 0xFFFFFFFF
    A = sc["albedo_np"].reshape(-1, 3)[hits[hit, 0]].astype(np.float64)
    r = I[hit, :3].astype(np.float64) - tgt[hit, :3]
    return float(np.sum(np.abs(np.sum(2 * r * A, axis=1))) * 2 / np.linalg.norm(S.SUN))


def oracle_shade(sc, albedo_np, sun=S.SUN):
    hits = sc["hits"].cpu().numpy().view(np.uint32)
    npix = hits.size // 3
    out = np.zeros((npix, 4), np.float32)
    O.lib().or_shade_lambert(hits.ctypes.data, npix, albedo_np.ctypes.data, O.fvec(sun),
                             AMBIENT, out.ctypes.data)
    return out


@pytest.mark.parametrize("traversal", [vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8])
def test_forward_fused_records_oracle(ctx, scene, traversal):
    import torch
    sc = scene
    W, H = sc["W"], sc["H"]
    f = vt.Frame.alloc(W, H)
    log(f"fused Lambert render traversal={traversal}")
    vt.render(ctx, sc["world"], sc["cam"], f,
              vt.params(W, H, 512, 512, S.SUN, vt.VT_FLAG_LAMBERT, traversal, AMBIENT,
                        sc["albedo"]))
    torch.cuda.synchronize()
    assert torch.equal(f.hits, sc["hits"])
    fused = f.shaded.reshape(-1, 4).cpu().numpy()
    rec = shade(ctx, sc, sc["albedo"]).cpu().numpy()
    ref = oracle_shade(sc, sc["albedo_np"])
    assert np.array_equal(rec.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(fused.view(np.uint32), ref.view(np.uint32))
    hit = sc["hits"].reshape(-1, 3)[:, 0].cpu().numpy().view(np.uint32) != 0xFFFFFFFF
    assert hit.sum() > 1000 and (~hit).sum() > 100  # both sky and terrain in view
    assert np.all(ref[~hit] == 0) and np.all(ref[hit, 3] == 1)


def test_oracle_render_lambert_agrees(ctx, scene):
    """The oracle's own fused Lambert render matches the GPU fused render."""
    import torch
    sc = scene
    W, H = sc["W"], sc["H"]
    f = vt.Frame.alloc(W, H)
    vt.render(ctx, sc["world"], sc["cam"], f,
              vt.params(W, H, 512, 512, S.SUN, vt.VT_FLAG_LAMBERT, 0, AMBIENT, sc["albedo"]))
    torch.cuda.synchronize()
    o = O.render(O.DenseWorld(sc["grid"], sc["pal"]), sc["cam"], W, H, 512, 512, S.SUN,
                 flags=vt.VT_FLAG_LAMBERT, ambient=AMBIENT, albedo=sc["albedo_np"],
                 gbuffer=False)
    assert np.array_equal(f.shaded.cpu().numpy().view(np.uint32), o["shaded"].view(np.uint32))


def test_compact_index_every_occupied_voxel(ctx, scene):
    import torch
    sc = scene
    grid = sc["grid"]
    n_occ = vt.build_compact_index(ctx, sc["world"])
    assert n_occ == int(np.count_nonzero(grid))
    assert vt.build_compact_index(ctx, sc["world"]) == n_occ  # idempotent
    occ, rank = compact_rank(grid)
    dense = sc["albedo_np"].reshape(-1, 3)
    comp = np.empty((n_occ, 3), np.float32)
    comp[rank] = dense[occ]
    # one record per occupied voxel, lit +x face: result depends only on albedo
    rec = np.zeros((n_occ, 3), np.uint32)
    rec[:, 0] = occ
    rec[:, 2] = 2
    rec_t = torch.from_numpy(rec.view(np.int32)).cuda()
    log(f"compact index over {n_occ} occupied voxels")
    a = shade(ctx, sc, sc["albedo"], hits=rec_t).cpu().numpy()
    b = shade(ctx, sc, torch.from_numpy(comp.ravel()).cuda(), flags=vt.VT_FLAG_ALBEDO_COMPACT,
              hits=rec_t).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # a planted wrong rank (off by one) would be seen
    comp_bad = np.roll(comp, 1, axis=0)
    c = shade(ctx, sc, torch.from_numpy(comp_bad.ravel()).cuda(),
              flags=vt.VT_FLAG_ALBEDO_COMPACT, hits=rec_t).cpu().numpy()
    assert not np.array_equal(a, c)


def test_backward_vs_fp64(ctx, scene):
    sc = scene
    hits = sc["hits"].cpu().numpy().view(np.uint32)
    npix = hits.size // 3
    tgt = sc["target"].cpu().numpy()
    log("backward L2 target vs FP64")
    ga, gs, loss = backward(ctx, sc, sc["albedo"], sc["target"], vt.VT_UPSTREAM_L2_TARGET)
    da = np.zeros(sc["albedo_np"].size, np.float64)
    ds = np.zeros(3, np.float64)
    loss64 = O.lib().or_lambert_grad64(hits.ctypes.data, npix, sc["albedo_np"].ctypes.data,
                                       O.fvec(S.SUN), AMBIENT, tgt.ctypes.data,
                                       da.ctypes.data, ds.ctypes.data)
    assert abs(loss - loss64) <= 2e-5 * abs(loss64)
    touched = da != 0
    assert touched.sum() > 1000
    np.testing.assert_allclose(ga, da, rtol=1e-5, atol=1e-5)
    scale = sun_term_bound(sc)
    log(f"grad_sun {gs} fp64 {ds} term bound {scale:.3g}")
    assert np.all(np.abs(gs - ds) <= 1e-5 * scale), (gs, ds, scale)
    assert np.linalg.norm(ds) > 1e-4 * scale  # the tolerance is < 10% of the gradient


def _l2(I, tgt):
    return float(np.sum((I[:, :3].astype(np.float64) - tgt[:, :3]) ** 2))


def test_finite_differences(ctx, scene):
    import torch
    sc = scene
    tgt = sc["target"].reshape(-1, 4).cpu().numpy().astype(np.float64)
    ga, gs, _ = backward(ctx, sc, sc["albedo"], sc["target"], vt.VT_UPSTREAM_L2_TARGET)
    vid = sc["hits"].reshape(-1, 3)[:, 0].cpu().numpy().view(np.uint32)
    visible = np.unique(vid[vid != 0xFFFFFFFF])
    rng = np.random.default_rng(7)
    picks = rng.choice(visible, 64, replace=False)
    h = 1e-3
    assert torch.is_tensor(sc["albedo"])
    log("albedo FD on 64 visible voxels")
    for v in picks:
        c = int(rng.integers(3))
        k = 3 * int(v) + c
        ap, am = sc["albedo"].clone(), sc["albedo"].clone()
        ap[k] += h
        am[k] -= h
        hp = float(ap[k].item()) - float(sc["albedo"][k].item())
        hm = float(sc["albedo"][k].item()) - float(am[k].item())
        lp = _l2(shade(ctx, sc, ap).cpu().numpy(), tgt)
        lm = _l2(shade(ctx, sc, am).cpu().numpy(), tgt)
        fd = (lp - lm) / (hp + hm)
        assert abs(fd - ga[k]) <= 1e-3 * max(1.0, abs(ga[k])), (v, c, fd, ga[k])
    log("sun FD")
    sun = np.array(S.SUN, np.float32)
    for k in range(3):
        sp, sm = sun.copy(), sun.copy()
        sp[k] += np.float32(h)
        sm[k] -= np.float32(h)
        lp = _l2(shade(ctx, sc, sc["albedo"], sun=sp).cpu().numpy(), tgt)
        lm = _l2(shade(ctx, sc, sc["albedo"], sun=sm).cpu().numpy(), tgt)
        fd = (lp - lm) / float(sp[k] - sm[k])
        assert abs(fd - gs[k]) <= 1e-3 * abs(gs[k]) + 0.05, (k, fd, gs[k])


@pytest.mark.parametrize("compact", [False, True])
def test_fused_grad_l2_matches_stored(ctx, scene, compact):
    import torch
    sc = scene
    W, H = sc["W"], sc["H"]
    flags = vt.VT_FLAG_LAMBERT | vt.VT_FLAG_GRAD_L2
    albedo = sc["albedo"]
    occ = rank = None
    if compact:
        n_occ = vt.build_compact_index(ctx, sc["world"])
        occ, rank = compact_rank(sc["grid"])
        comp = np.empty((n_occ, 3), np.float32)
        comp[rank] = sc["albedo_np"].reshape(-1, 3)[occ]
        albedo = torch.from_numpy(comp.ravel()).cuda()
        flags |= vt.VT_FLAG_ALBEDO_COMPACT
    ga = torch.zeros_like(albedo)
    gs = torch.zeros(3, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    f = vt.Frame.alloc(W, H)
    f.target = sc["target"]
    log(f"fused GRAD_L2 render compact={compact}")
    vt.render(ctx, sc["world"], sc["cam"], f,
              vt.params(W, H, 512, 512, S.SUN, flags, 0, AMBIENT, albedo, ga, gs, loss))
    torch.cuda.synchronize()
    assert torch.equal(f.hits, sc["hits"])
    ref = oracle_shade(sc, sc["albedo_np"])
    assert np.array_equal(f.shaded.reshape(-1, 4).cpu().numpy().view(np.uint32),
                          ref.view(np.uint32))
    ga2, gs2, loss2 = backward(ctx, sc, sc["albedo"], sc["target"], vt.VT_UPSTREAM_L2_TARGET)
    g = ga.cpu().numpy()
    if compact:
        dense = np.zeros_like(ga2).reshape(-1, 3)
        dense[occ] = g.reshape(-1, 3)[rank]
        g = dense.ravel()
    np.testing.assert_allclose(g, ga2, rtol=1e-5, atol=1e-5)
    assert abs(loss.item() - loss2) <= 2e-5 * abs(loss2)
    assert np.all(np.abs(gs.cpu().numpy() - gs2) <= 1e-5 * sun_term_bound(sc))


def test_autograd_upstream_kind(ctx, scene):
    import torch
    sc = scene

    class LambertShade(torch.autograd.Function):
        @staticmethod
        def forward(fctx, albedo, sun):
            fctx.save_for_backward(albedo, sun)
            return shade(ctx, sc, albedo, sun=sun.tolist())

        @staticmethod
        def backward(fctx, gout):
            albedo, sun = fctx.saved_tensors
            ga, gs, _ = backward(ctx, sc, albedo, gout.contiguous(), vt.VT_UPSTREAM_GRADIENT,
                                 sun=sun.tolist())
            return torch.from_numpy(ga).cuda(), torch.from_numpy(gs).cuda()

    albedo = sc["albedo"].clone().requires_grad_(True)
    sun = torch.tensor(S.SUN, dtype=torch.float32, device="cuda", requires_grad=True)
    I = LambertShade.apply(albedo, sun)
    loss = ((I[:, :3] - sc["target"].reshape(-1, 4)[:, :3]) ** 2).sum()
    loss.backward()
    ga2, gs2, loss2 = backward(ctx, sc, sc["albedo"], sc["target"], vt.VT_UPSTREAM_L2_TARGET)
    np.testing.assert_allclose(albedo.grad.cpu().numpy(), ga2, rtol=1e-5, atol=1e-5)
    assert np.all(np.abs(sun.grad.cpu().numpy() - gs2) <= 1e-5 * sun_term_bound(sc))
    assert abs(loss.item() - loss2) <= 2e-5 * abs(loss2)


def test_error_paths(ctx, scene):
    import torch
    sc = scene
    W, H = sc["W"], sc["H"]
    f = vt.Frame.alloc(W, H)
    ga = torch.zeros_like(sc["albedo"])
    with pytest.raises(vt.VtError) as e:  # GRAD_L2 without a target image
        vt.render(ctx, sc["world"], sc["cam"], f,
                  vt.params(W, H, 512, 512, S.SUN, vt.VT_FLAG_LAMBERT | vt.VT_FLAG_GRAD_L2, 0,
                            AMBIENT, sc["albedo"], ga))
    assert e.value.status == N.VT_ERR_NULL_POINTER
    with pytest.raises(vt.VtError) as e:  # GRAD_L2 without LAMBERT
        vt.render(ctx, sc["world"], sc["cam"], f,
                  vt.params(W, H, 512, 512, S.SUN, vt.VT_FLAG_GRAD_L2, 0, AMBIENT,
                            sc["albedo"], ga))
    assert e.value.status == N.VT_ERR_INVALID_ARGUMENT
    with pytest.raises(vt.VtError) as e:  # unknown upstream kind
        backward(ctx, sc, sc["albedo"], sc["target"], 7)
    assert e.value.status == N.VT_ERR_INVALID_ARGUMENT
    with pytest.raises(vt.VtError) as e:  # no albedo
        prm = vt.params(1, 1, flags=vt.VT_FLAG_LAMBERT)
        out = torch.empty((4, 4), device="cuda")
        vt.lambert_shade(ctx, sc["world"], sc["hits"].reshape(-1, 3)[:4].contiguous(), prm, out)
    assert e.value.status == N.VT_ERR_NULL_POINTER


def test_zero_sun_rejected_for_lambert(ctx, scene):
    """Lambert shading divides by |sun|: a sun of zero length (or one whose
    squared length underflows) is VT_ERR_INVALID_ARGUMENT, not NaN output;
    the reference composite keeps accepting it (its 0.001 substitution)."""
    import torch
    sc = scene
    W, H = sc["W"], sc["H"]
    f = vt.Frame.alloc(W, H)
    for sun in ((0.0, 0.0, 0.0), (1e-30, 0.0, -1e-30)):
        with pytest.raises(vt.VtError) as e:
            vt.render(ctx, sc["world"], sc["cam"], f,
                      vt.params(W, H, 512, 512, sun, vt.VT_FLAG_LAMBERT, 0, AMBIENT, sc["albedo"]))
        assert e.value.status == N.VT_ERR_INVALID_ARGUMENT
        with pytest.raises(vt.VtError) as e:
            shade(ctx, sc, sc["albedo"], sun=sun)
        assert e.value.status == N.VT_ERR_INVALID_ARGUMENT
        with pytest.raises(vt.VtError) as e:
            backward(ctx, sc, sc["albedo"], sc["target"], vt.VT_UPSTREAM_L2_TARGET, sun=sun)
        assert e.value.status == N.VT_ERR_INVALID_ARGUMENT
    vt.render(ctx, sc["world"], sc["cam"], f, vt.params(W, H, 512, 512, (0.0, 0.0, 0.0)))
    torch.cuda.synchronize()
    assert torch.isfinite(f.shaded).all()
