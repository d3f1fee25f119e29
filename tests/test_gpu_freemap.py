"""The free-region test on the GPU: a ray standing in a 16^3 region from
which every region ahead of it (on all three axes, in its travel direction)
is empty ends as a miss.  Every output but the step counts must equal the
oracle's, which runs every ray to its end.

Worlds with large empty parts (a few boxes and single voxels in a larger
grid, at 16^3 .. 256^3: 1 to 16 regions per axis), cameras inside and
outside the grid looking into all eight direction octants, suns of either
sign; the map kept current by edits on both layouts (a stale map would end
rays before voxels added after it was built); the flag rules.
"""
import numpy as np
import pytest

import oracle_ffi as O
import vt
from test_gpu_parity import compare, poison
from vt import camera as vcam
from vt import scenes as S

pytestmark = pytest.mark.gpu

ON = vt.VT_FLAG_FREE_MAP_ON
OFF = vt.VT_FLAG_FREE_MAP_OFF


def log(msg):
    print(f"[freemap] {msg}", flush=True)


def boxes_world(n, seed, n_boxes=6, n_single=12):
    """A dim n grid holding a few boxes and single voxels: most 16^3 regions
    empty, and empty orthants in every octant from many of them."""
    rng = np.random.default_rng(seed)
    g = np.zeros((n, n, n), np.uint8)
    for _ in range(n_boxes):
        s = rng.integers(1, max(2, n // 6), 3)
        o = rng.integers(0, n - s + 1)
        g[o[2]:o[2] + s[2], o[1]:o[1] + s[1], o[0]:o[0] + s[0]] = rng.integers(1, 9)
    for _ in range(n_single):
        x, y, z = rng.integers(0, n, 3)
        g[z, y, x] = rng.integers(1, 9)
    return g


def render_free(ctx, world, cam, W, H, cap_p, cap_s, sun, flags=ON, gbuffer=True, traversal=0):
    import torch
    f = poison(vt.Frame.alloc(W, H, gbuffer=gbuffer))
    vt.render(ctx, world, cam, f, vt.params(W, H, cap_p, cap_s, sun, flags, traversal))
    torch.cuda.synchronize()
    out = {"shaded": f.shaded.cpu().numpy(), "hits": f.hits.cpu().numpy().view(np.uint32)}
    if gbuffer:
        out.update(color=f.color.cpu().numpy(), normal=f.normal.cpu().numpy(),
                   position=f.position.cpu().numpy(),
                   illumination=f.illumination.cpu().numpy())
    return out


def random_cameras(n, rng, count):
    """Eyes inside and outside the grid, targets anywhere: every octant."""
    out = []
    for _ in range(count):
        inside = rng.random() < 0.5
        lo, hi = (0.0, float(n)) if inside else (-0.5 * n, 1.5 * n)
        eye = rng.uniform(lo, hi, 3)
        tgt = rng.uniform(-0.5 * n, 1.5 * n, 3)
        if np.linalg.norm(tgt - eye) < 1.0:
            tgt = eye + 7.0
        out.append(vcam.look_at(tuple(eye), tuple(tgt), float(rng.uniform(40, 100))))
    return out


SUNS = [S.SUN, (-0.5, 0.7, -0.2), (0.3, -0.8, 0.4), (-0.6, -0.5, -0.6)]


@pytest.mark.parametrize("n", [16, 32, 64, 256])
def test_random_boxes_every_octant(ctx, n):
    import brickmap_util as BU
    pal = S.palette()
    rng = np.random.default_rng(1000 + n)
    grid = boxes_world(n, 7 + n)
    wd = vt.World.dense(ctx, grid, pal)
    dim, chunks, pool, atlas = BU.embedding_a(grid, pal)
    wb = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    ow = O.DenseWorld(grid, pal)
    W, H = 48, 40
    cap = 4 * n
    for k, cam in enumerate(random_cameras(n, rng, 12)):
        sun = SUNS[k % len(SUNS)]
        o = O.render(ow, cam, W, H, cap, cap, sun)
        label = f"boxes n={n} camera {k} sun={sun}"
        log(label)
        compare(render_free(ctx, wd, cam, W, H, cap, cap, sun), o, stats=False,
                label=label + " dense")
        gb = render_free(ctx, wb, cam, W, H, cap, cap, sun, gbuffer=False)
        assert np.array_equal(gb["hits"], o["hits"]), f"{label} reference layout: hits"
        assert np.max(np.abs(gb["shaded"] - o["shaded"])) <= 1e-5, f"{label} reference layout"
    wd.close()
    wb.close()


def test_map_follows_dense_edits(ctx):
    """A box added to an empty world after the map was built is hit; removed
    again, every ray misses."""
    pal = S.palette()
    n = 64
    grid = np.zeros((n, n, n), np.uint8)
    w = vt.World.dense(ctx, grid, pal)
    cam = vcam.look_at((-20.0, 70.0, -25.0), (32.0, 30.0, 32.0), 60.0)
    W, H = 64, 48
    o = O.render(O.DenseWorld(grid, pal), cam, W, H, 512, 512, S.SUN)
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False, label="empty")
    assert np.all(o["hits"][..., 0] == 0xFFFFFFFF)
    z, y, x = np.mgrid[20:44, 10:30, 24:40]
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.uint32)
    w.set_voxels(xyz, np.full(len(xyz), 5, np.uint32))
    grid[20:44, 10:30, 24:40] = 5
    o = O.render(O.DenseWorld(grid, pal), cam, W, H, 512, 512, S.SUN)
    assert np.mean(o["hits"][..., 0] != 0xFFFFFFFF) > 0.05
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False, label="box added")
    w.set_voxels(xyz, np.zeros(len(xyz), np.uint32))
    grid[:] = 0
    o = O.render(O.DenseWorld(grid, pal), cam, W, H, 512, 512, S.SUN)
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False, label="box removed")
    w.close()


def test_map_follows_brickmap_edits(ctx):
    """An empty reference-layout world built up by an engine-style edit stream
    (the 29 model slots of the engine's assets) in three batches: after each,
    the test on gives exactly the images of the test off."""
    from test_gpu_edit import atlas_fixture, engine_edit_stream
    n = 64
    w = vt.World.empty_brickmap(ctx, n, atlas_fixture())
    cam = vcam.look_at((-20.0, 40.0, -25.0), (32.0, 6.0, 32.0), 60.0)
    W, H = 64, 48
    before = render_free(ctx, w, cam, W, H, 1536, 512, S.SUN, gbuffer=False)
    assert np.all(before["hits"][..., 0] == 0xFFFFFFFF)
    ed = engine_edit_stream(n, 9)
    for k, part in enumerate(np.array_split(ed, 3)):
        w.set_voxels(part[:, :3], part[:, 3])
        on = render_free(ctx, w, cam, W, H, 1536, 512, S.SUN, gbuffer=False)
        off = render_free(ctx, w, cam, W, H, 1536, 512, S.SUN, flags=OFF, gbuffer=False)
        log(f"brickmap edit batch {k}: {np.mean(off['hits'][..., 0] != 0xFFFFFFFF):.2f} hit")
        assert np.array_equal(on["hits"], off["hits"]), f"batch {k}: hits"
        assert np.array_equal(on["shaded"].view(np.uint32), off["shaded"].view(np.uint32))
    assert np.mean(off["hits"][..., 0] != 0xFFFFFFFF) > 0.3
    w.close()


def test_on_off_auto_agree_c1_row_band(ctx):
    """The headline world: forced on, forced off and the default give the
    same images (the default takes the test: caps of 2048)."""
    import torch
    grid = S.generate(S.spec("c1"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    del grid
    cam, W, H = S.views("c1")[0]
    imgs = []
    for flags in (0, ON, OFF):
        f = vt.Frame.alloc(W, H)
        vt.render(ctx, w, cam, f, vt.params(W, H, 2048, 2048, S.SUN, flags))
        torch.cuda.synchronize()
        imgs.append((f.hits.cpu().numpy().view(np.uint32), f.shaded.cpu().numpy().view(np.uint32)))
    for k in (1, 2):
        assert np.array_equal(imgs[0][0], imgs[k][0])
        assert np.array_equal(imgs[0][1], imgs[k][1])
    # which launches take the test (vt_free_map_used)
    p = vt.params(W, H, 2048, 2048, S.SUN)
    assert vt.free_map_used(w, vt.Frame.alloc(W, H), p)
    assert not vt.free_map_used(w, vt.Frame.alloc(W, H, steps=True), p)
    assert not vt.free_map_used(w, vt.Frame.alloc(W, H, stats=True), p)
    assert vt.free_map_used(w, vt.Frame.alloc(W, H, stats=True),
                            vt.params(W, H, 2048, 2048, S.SUN, ON))
    assert not vt.free_map_used(w, vt.Frame.alloc(W, H), vt.params(W, H, 2048, 2048, S.SUN, OFF))
    assert not vt.free_map_used(w, vt.Frame.alloc(W, H), vt.params(W, H, 192, 48, S.SUN))
    assert not vt.free_map_used(w, vt.Frame.alloc(W, H),
                                vt.params(W, H, 2048, 2048, S.SUN, 0, vt.VT_TRAVERSAL_BITMASK))
    w.close()
    # a world with free regions in no octant: the default leaves the test off
    g = boxes_world(64, 3)
    g[:, :, :] |= (np.arange(64)[None, None, :] % 16 == 15).astype(np.uint8)  # a wall per region
    w = vt.World.dense(ctx, g, pal)
    assert not vt.free_map_used(w, vt.Frame.alloc(64, 64), vt.params(64, 64, 2048, 2048, S.SUN))
    w.close()


def _morton3(x, y, z):
    def part(t):
        t = np.asarray(t, np.uint64) & 0x3FF
        t = (t | (t << 16)) & 0xFF0000FF
        t = (t | (t << 8)) & 0x0300F00F
        t = (t | (t << 4)) & 0x030C30C3
        return (t | (t << 2)) & 0x09249249
    return part(x) | (part(y) << 1) | (part(z) << 2)


def _region_occupancy(grid):
    """[z, y, x] regions of 16^3 blocks (one region for smaller grids) that
    hold a nonzero block."""
    n = grid.shape[0]
    if n < 16:
        return np.array([[[grid.any()]]])
    R = n // 16
    return grid.reshape(R, 16, R, 16, R, 16).any(axis=(1, 3, 5))


def _pack(free):
    """Octant bitmaps as the library lays them out: bit r = region with
    Morton code r (free is [8, z, y, x] booleans)."""
    R = free.shape[1]
    z, y, x = np.meshgrid(np.arange(R), np.arange(R), np.arange(R), indexing="ij")
    codes = _morton3(x.ravel(), y.ravel(), z.ravel()).astype(np.int64)
    words = max(1, (R ** 3 + 31) // 32)
    out = np.zeros((8, words), np.uint32)
    for o in range(8):
        bits = free[o].ravel()
        for c, b in zip(codes, bits):
            if b:
                out[o, c >> 5] |= np.uint32(1 << (c & 31))
    return out


def expected_free_brute(grid):
    """The definition, region pair by region pair: region r is free for
    octant o when no occupied region r' lies at or ahead of r on every axis
    in o's travel direction."""
    occ = _region_occupancy(grid)
    R = occ.shape[0]
    occupied = np.argwhere(occ)  # [z, y, x]
    free = np.ones((8, R, R, R), bool)
    for o in range(8):
        sx, sy, sz = (-1 if o & 1 else 1), (-1 if o & 2 else 1), (-1 if o & 4 else 1)
        for z in range(R):
            for y in range(R):
                for x in range(R):
                    for (oz, oy, ox) in occupied:
                        if (ox - x) * sx >= 0 and (oy - y) * sy >= 0 and (oz - z) * sz >= 0:
                            free[o, z, y, x] = False
                            break
    return _pack(free)


def expected_free_scan(grid):
    """The same map from running ORs of the region occupancy along each axis
    in the travel direction (for grids too large for the pairwise form)."""
    occ = _region_occupancy(grid)
    free = np.zeros((8,) + occ.shape, bool)
    for o in range(8):
        a = occ.copy()
        for axis, bit in ((2, 0), (1, 1), (0, 2)):
            if (o >> bit) & 1:
                a = np.maximum.accumulate(a, axis=axis)
            else:
                a = np.flip(np.maximum.accumulate(np.flip(a, axis), axis=axis), axis)
        free[o] = ~a
    return _pack(free)


@pytest.mark.parametrize("n", [16, 32, 64, 128])
def test_map_equals_its_definition(ctx, n):
    """The device-built map, bit for bit, against the definition evaluated
    region pair by region pair, on the dense layout and (same occupancy) the
    reference layout, then again after edits that add and clear blocks."""
    import brickmap_util as BU
    pal = S.palette()
    grid = boxes_world(n, 500 + n, n_boxes=3, n_single=4)
    want = expected_free_brute(grid)
    wd = vt.World.dense(ctx, grid, pal)
    got, valid = wd.read_free_map()
    assert valid and np.array_equal(got, want), f"dense n={n}"
    dim, chunks, pool, atlas = BU.embedding_a(grid, pal)
    wb = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    got, valid = wb.read_free_map()
    assert valid and np.array_equal(got, want), f"reference layout n={n}"
    wb.close()
    # edits: a voxel in a far corner region and the removal of a box's corner
    rng = np.random.default_rng(n)
    xyz = rng.integers(0, n, (6, 3)).astype(np.uint32)
    vals = np.array([3, 0, 5, 0, 7, 2], np.uint32)
    wd.set_voxels(xyz, vals)
    for (x, y, z), v in zip(xyz, vals):
        grid[z, y, x] = v
    got, valid = wd.read_free_map()
    assert valid and np.array_equal(got, expected_free_brute(grid)), f"dense n={n} after edits"
    wd.close()


def test_map_of_empty_and_solid_worlds(ctx):
    """A world with nothing in it is free everywhere for every octant; one
    with a block in every region nowhere; the smallest reference layout (8^3
    blocks, one region) too."""
    pal = S.palette()
    for n, fill, want_bit in ((16, 0, 1), (64, 0, 1), (64, 9, 0)):
        g = np.zeros((n, n, n), np.uint8)
        if fill:
            g[::16, ::16, ::16] = 2  # one block per region
        w = vt.World.dense(ctx, g, pal)
        got, valid = w.read_free_map()
        R = max(n // 16, 1)
        full = np.zeros(got.shape[1], np.uint32)
        for r in range(R ** 3):
            full[r >> 5] |= np.uint32((1 << (r & 31)) if want_bit else 0)
        assert valid and all(np.array_equal(got[o], full) for o in range(8)), (n, fill)
        w.close()
    atlas = np.zeros((256, 256, 256), np.uint32)
    w = vt.World.brickmap(ctx, 8, np.zeros(1, np.uint32), np.zeros(0, np.uint32), atlas)
    got, valid = w.read_free_map()
    assert valid and got.shape == (8, 1) and np.all(got == 1)
    w.close()


def test_map_of_the_headline_world(ctx):
    """The 512^3 headline world (32^3 regions): the device map against the
    running-OR form; 84% of the regions free for up-going rays."""
    grid = S.generate(S.spec("c1"))
    w = vt.World.dense(ctx, grid, S.palette())
    got, valid = w.read_free_map()
    want = expected_free_scan(grid)
    assert valid and np.array_equal(got, want)
    share = [sum(bin(int(x)).count("1") for x in got[o]) / 32 ** 3 for o in range(8)]
    log(f"c1 share of free regions per octant: {[round(x, 3) for x in share]}")
    assert max(share) > 0.8
    w.close()


def test_statistics_of_executed_steps(ctx):
    """Statistics with the test forced on count the steps the kernel
    executed: fewer than traceMap's (rays end early), with the same hits,
    occluded shadow rays, material reads and entity hits; the images equal."""
    import torch
    n = 256
    grid = boxes_world(n, 99, n_boxes=10, n_single=40)
    w = vt.World.dense(ctx, grid, S.palette())
    cam = vcam.look_at((-40.0, 180.0, -60.0), (128.0, 40.0, 128.0), 70.0)
    W, H = 320, 240
    out = {}
    for name, flags in (("ref", 0), ("on", ON)):
        f = poison(vt.Frame.alloc(W, H, stats=True))
        vt.render(ctx, w, cam, f, vt.params(W, H, 1024, 1024, S.SUN, flags))
        torch.cuda.synchronize()
        out[name] = (f.stats_dict(), f.hits.cpu().numpy().view(np.uint32),
                     f.shaded.cpu().numpy().view(np.uint32))
    ref, on = out["ref"][0], out["on"][0]
    log(f"steps executed {on['primary_steps']} + {on['shadow_steps']} of traceMap's "
        f"{ref['primary_steps']} + {ref['shadow_steps']}")
    for k in ("primary_hits", "occluded"):
        assert on[k] == ref[k], k
    assert on["primary_steps"] < ref["primary_steps"]
    assert on["shadow_steps"] <= ref["shadow_steps"]
    assert np.array_equal(out["ref"][1], out["on"][1])
    assert np.array_equal(out["ref"][2], out["on"][2])
    w.close()


def test_flag_rules(ctx):
    grid = np.zeros((16, 16, 16), np.uint8)
    w = vt.World.dense(ctx, grid, S.palette())
    f = vt.Frame.alloc(8, 8)
    cam = vcam.look_at((8.0, 8.0, -10.0), (8.0, 8.0, 8.0), 60.0)
    with pytest.raises(vt.VtError) as e:
        vt.render(ctx, w, cam, f, vt.params(8, 8, 64, 64, S.SUN, ON | OFF))
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    w.close()


def grazing_cameras(n, y_top, rng, count):
    """Eyes just above (or below) a slab's face, looking across it at small
    pitches in every compass direction: rays that pass the box of the nonempty
    voxels by fractions of a voxel, hit it far away, or leave the grid."""
    out = []
    for k in range(count):
        eye = np.array([rng.uniform(0.0, n), y_top + rng.uniform(0.2, 3.0), rng.uniform(0.0, n)])
        if k % 3 == 2:  # from outside the grid
            eye[[0, 2]] = rng.choice([-0.3 * n, 1.3 * n], 2)
        # toward the middle of the grid, give or take 60 degrees
        yaw = np.arctan2(0.5 * n - eye[2], 0.5 * n - eye[0]) + rng.uniform(-1.05, 1.05)
        pitch = rng.uniform(-0.05, 0.05)
        tgt = eye + np.array([np.cos(yaw), np.tan(pitch), np.sin(yaw)]) * n
        out.append(vcam.look_at(tuple(eye), tuple(tgt), float(rng.uniform(20, 60))))
    return out


@pytest.mark.parametrize("n", [64, 256])
def test_reach_box_grazing_rays(ctx, n):
    """Primary rays whose line misses the box of the nonempty voxels
    (widened by the traversal's reach) end as misses before the walk: a slab
    seen from just above its top face at grazing angles, so that many lines
    pass the box by less than the widening and others by more.  Every output
    but the step counts equals the oracle's, with the free-region test and
    without it (the box is taken whenever a frame stores no step counts);
    the reference layout of the same voxels (its box scanned from the block
    pool) gives the same hits."""
    import brickmap_util as BU
    pal = S.palette()
    rng = np.random.default_rng(77 + n)
    g = np.zeros((n, n, n), np.uint8)
    y0, y1 = n // 4, n // 4 + 3
    g[n // 8:7 * n // 8, y0:y1, n // 8:7 * n // 8] = 4
    g[n // 2, y1:y1 + n // 8, n // 2] = 6  # a post on it
    wd = vt.World.dense(ctx, g, pal)
    wb = vt.World.brickmap(ctx, *BU.embedding_a(g, pal))
    ow = O.DenseWorld(g, pal)
    W, H = 64, 48
    cap = 4 * n
    hit_share = []
    for k, cam in enumerate(grazing_cameras(n, float(y1), rng, 9)):
        sun = SUNS[k % len(SUNS)]
        o = O.render(ow, cam, W, H, cap, cap, sun)
        label = f"grazing n={n} camera {k}"
        log(label)
        compare(render_free(ctx, wd, cam, W, H, cap, cap, sun), o, stats=False, label=label)
        # the box without the free-region test (frames that store no step
        # counts), flat and bitmask traversals
        for trav in (vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK):
            compare(render_free(ctx, wd, cam, W, H, cap, cap, sun, flags=OFF, traversal=trav), o,
                    stats=False, label=f"{label} test off, traversal {trav}")
        gb = render_free(ctx, wb, cam, W, H, cap, cap, sun, gbuffer=False)
        assert np.array_equal(gb["hits"], o["hits"]), f"{label} reference layout: hits"
        assert np.max(np.abs(gb["shaded"] - o["shaded"])) <= 1e-5, f"{label} reference layout"
        hit_share.append(float(np.mean(o["hits"][..., 0] != 0xFFFFFFFF)))
    log(f"grazing n={n}: hit shares {[round(x, 2) for x in hit_share]}")
    assert 0.05 < np.mean(hit_share) < 0.95
    wd.close()
    wb.close()


def test_reach_box_grows_with_edits(ctx):
    """The box follows edits: voxels set far outside it after the world was
    made are hit (rays toward them are not ended by the stale box), on both
    layouts, and a world made empty and filled again by edits renders like
    the oracle."""
    pal = S.palette()
    n = 128
    g = np.zeros((n, n, n), np.uint8)
    g[10:20, 2:5, 10:20] = 3
    w = vt.World.dense(ctx, g, pal)
    cam = vcam.look_at((64.0, 70.0, -40.0), (64.0, 90.0, 64.0), 50.0)
    W, H = 64, 48
    o = O.render(O.DenseWorld(g, pal), cam, W, H, 512, 512, S.SUN)
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False, label="small box")
    # a wall high above the box, in front of the camera
    z, y, x = np.mgrid[80:82, 80:110, 40:90]
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.uint32)
    w.set_voxels(xyz, np.full(len(xyz), 7, np.uint32))
    g[80:82, 80:110, 40:90] = 7
    o = O.render(O.DenseWorld(g, pal), cam, W, H, 512, 512, S.SUN)
    assert np.mean(o["hits"][..., 0] != 0xFFFFFFFF) > 0.1
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False, label="wall added")
    w.close()
    # an empty world (no box: nothing ends early) filled by edits
    g = np.zeros((n, n, n), np.uint8)
    w = vt.World.dense(ctx, g, pal)
    o = O.render(O.DenseWorld(g, pal), cam, W, H, 512, 512, S.SUN)
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False, label="empty")
    w.set_voxels(xyz, np.full(len(xyz), 7, np.uint32))
    g[80:82, 80:110, 40:90] = 7
    o = O.render(O.DenseWorld(g, pal), cam, W, H, 512, 512, S.SUN)
    compare(render_free(ctx, w, cam, W, H, 512, 512, S.SUN), o, stats=False,
            label="empty world, wall added")
    w.close()
    # the reference layout: the small box's world, then the wall as block edits
    import brickmap_util as BU
    g = np.zeros((n, n, n), np.uint8)
    g[10:20, 2:5, 10:20] = 3
    wb = vt.World.brickmap(ctx, *BU.embedding_a(g, pal))
    wb.set_voxels(xyz, np.full(len(xyz), 7, np.uint32))
    g[80:82, 80:110, 40:90] = 7
    o = O.render(O.DenseWorld(g, pal), cam, W, H, 512, 512, S.SUN)
    gb = render_free(ctx, wb, cam, W, H, 512, 512, S.SUN, gbuffer=False)
    assert np.array_equal(gb["hits"], o["hits"]), "reference layout, wall added: hits"
    assert np.max(np.abs(gb["shaded"] - o["shaded"])) <= 1e-5
    wb.close()


def test_blind_steps_counted(ctx):
    """A camera above the box of the nonempty voxels: primary rays take their
    first steps without lookups.  The statistics of the steps executed
    (VT_FLAG_FREE_MAP_ON) count them in the l1_fetches slot, fewer than the
    primary steps; the images equal traceMap's (statistics of traceMap's
    steps count none)."""
    import torch
    n = 128
    g = np.zeros((n, n, n), np.uint8)
    g[:, 0:20, :] = 3
    g[40:60, 20:30, 50:70] = 5
    w = vt.World.dense(ctx, g, S.palette())
    cam = vcam.look_at((64.0, 110.0, -30.0), (64.0, 10.0, 100.0), 60.0)
    W, H = 96, 64
    out = {}
    for name, flags in (("ref", 0), ("on", ON)):
        f = poison(vt.Frame.alloc(W, H, stats=True))
        vt.render(ctx, w, cam, f, vt.params(W, H, 1024, 1024, S.SUN, flags))
        torch.cuda.synchronize()
        out[name] = (f.stats_dict(), f.hits.cpu().numpy().view(np.uint32),
                     f.shaded.cpu().numpy().view(np.uint32))
    ref, on = out["ref"][0], out["on"][0]
    log(f"blind steps {on['l1_fetches']} of {on['primary_steps']} primary steps executed "
        f"(traceMap's {ref['primary_steps']})")
    assert ref["l1_fetches"] == 0
    assert 0 < on["l1_fetches"] < on["primary_steps"] <= ref["primary_steps"]
    assert np.array_equal(out["ref"][1], out["on"][1])
    assert np.array_equal(out["ref"][2], out["on"][2])
    o = O.render(O.DenseWorld(g, S.palette()), cam, W, H, 1024, 1024, S.SUN)
    assert np.array_equal(out["on"][1], o["hits"])
    w.close()
