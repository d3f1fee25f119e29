"""World edits, voxel reads and the centre-ray pick on the GPU, through the C-ABI.

* Brickmap worlds built by edits: an engine-style edit stream (water base
  layer, terrain columns of the reference's model slots with overwrites,
  grass and flower blocks without the solid bit, trees) applied in batches to
  an empty map must read back exactly like the CPU restatement of
  VoxelBrickmap.set/get (oracle/voxel_map.py), grow the pool, and render --
  through the model atlas of the engine's 29 slots (tests/golden) -- exactly
  like the oracle renders the chunk table and pool that the restatement
  built.
* Dense worlds: edits that clear and add voxels (with repeats) render like
  the oracle's render of the edited grid, for both traversals (the bitmask
  traversal reads the rebuilt occupancy words), and keep the compact index
  equal to the occupied count.
* vt_pick: every field equal to the oracle's or_pick.
"""
import os
import sys

import numpy as np
import pytest

import oracle_ffi as O
import vt
from vt import camera as vcam
from vt import scenes as S

sys.path.insert(0, O.ORACLE_DIR)
import voxel_map as VM  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "atlas_assets.npz")


def log(msg):
    print(f"[edit] {msg}", flush=True)


def atlas_fixture():
    g = np.load(GOLDEN)
    t = np.zeros(256 ** 3, np.uint32)
    t[g["index"]] = g["value"]
    return t


def engine_edit_stream(dim, seed):
    """Edits in the style of procgen.zig:6-70 over block ids of the 29 atlas
    slots: (x, y, z, value) in order, with repeated positions."""
    rng = np.random.default_rng(seed)
    ed = []
    for x in range(dim):
        for z in range(dim):
            for y in range(3):
                ed.append((x, y, z, VM.voxel(13)))  # water base layer
    for x in range(dim):
        for z in range(dim):
            h = int(6 + 3 * np.sin(x / 6.0) + 2 * np.cos(z / 4.0) + rng.integers(0, 2))
            for y in range(h):
                ed.append((x, y, z, VM.voxel(21 + int(rng.integers(3)))))  # dirt
                if y <= 4:
                    ed.append((x, y, z, VM.voxel(25 + int(rng.integers(3)))))  # sand
                elif y == h - 1:
                    ed.append((x, y, z, VM.voxel(int(rng.integers(6)))))  # surface top
            if h + 1 < dim and rng.integers(5) == 0:
                ed.append((x, h, z, VM.voxel(7 + int(rng.integers(5)), solid=False)))  # grass
            if h + 1 < dim and rng.integers(41) == 0:
                ed.append((x, h, z, VM.voxel(12)))  # flower
            if 2 < x < dim - 4 and 2 < z < dim - 4 and h + 10 < dim and rng.integers(120) == 0:
                th = 4 + int(rng.integers(4))
                for o in range(th):
                    ed.append((x + 1, h + o, z + 1, VM.voxel(14 + int(rng.integers(3)))))
                for a in range(3):
                    for b in range(3):
                        for c in range(3):
                            ed.append((x + a, h + th + b, z + c,
                                       VM.voxel(18 + int(rng.integers(2)))))
    return np.array(ed, dtype=np.uint32)


@pytest.fixture(scope="module")
def edited_brickmap(ctx):
    dim = 64
    atlas = atlas_fixture()
    ed = engine_edit_stream(dim, 5)
    ref = VM.VoxelBrickmap(dim)
    for x, y, z, v in ed:
        ref.set(int(x), int(y), int(z), int(v))
    w = vt.World.empty_brickmap(ctx, dim, atlas)
    b0 = w.device_bytes()
    log(f"{len(ed)} edits in 3 batches")
    for part in np.array_split(ed, 3):
        w.set_voxels(part[:, :3], part[:, 3])
    yield dict(dim=dim, atlas=atlas, ref=ref, world=w, grew=w.device_bytes() - b0, edits=ed)
    w.close()


def test_brickmap_edits_read_back_like_voxelbrickmap(ctx, edited_brickmap):
    e = edited_brickmap
    dim, ref, w = e["dim"], e["ref"], e["world"]
    g = np.stack(np.meshgrid(np.arange(dim), np.arange(dim), np.arange(dim), indexing="ij"),
                 -1).reshape(-1, 3)
    got = w.get_voxels(g)
    exp = np.array([ref.get(int(x), int(y), int(z)) for x, y, z in g], np.uint32)
    assert np.array_equal(got, exp)
    assert e["grew"] >= (len(ref.blocks) - 1) * 2048  # pool grew past the initial block
    assert np.count_nonzero(exp) > dim * dim * 4


def test_brickmap_edited_world_renders_like_oracle(ctx, edited_brickmap):
    import torch
    e = edited_brickmap
    dim, ref, w = e["dim"], e["ref"], e["world"]
    ow = O.BrickmapWorld(dim, ref.chunks, ref.pool(), e["atlas"])
    for k, (eye, tgt) in enumerate([((32.0, 30.0, -12.0), (32.0, 6.0, 32.0)),
                                    ((-8.0, 22.0, 70.0), (40.0, 4.0, 20.0))]):
        cam = vcam.look_at(eye, tgt, 70.0)
        W, H = 160, 120
        f = vt.Frame.alloc(W, H, gbuffer=True, steps=True)
        log(f"brickmap render view {k}")
        vt.render(ctx, w, cam, f, vt.params(W, H, 1536, 512, S.SUN))
        torch.cuda.synchronize()
        o = O.render(ow, cam, W, H, 1536, 512, S.SUN)
        gh = f.hits.cpu().numpy().view(np.uint32)
        assert np.array_equal(gh, o["hits"]), f"view {k}: {np.sum(np.any(gh != o['hits'], -1))}"
        assert np.array_equal(f.steps.cpu().numpy().view(np.uint32), o["steps"])
        assert np.max(np.abs(f.shaded.cpu().numpy() - o["shaded"])) <= 1e-5
        hit = gh[..., 0] != 0xFFFFFFFF
        assert hit.mean() > 0.3
        # model texels were hit (not only solid cubes): sub-voxel ids vary
        assert len(np.unique((gh[hit, 2] >> 4) & 511)) > 50


def test_pick_brickmap_and_dense(ctx, edited_brickmap):
    e = edited_brickmap
    ow = O.BrickmapWorld(e["dim"], e["ref"].chunks, e["ref"].pool(), e["atlas"])
    grid = S.generate(S.spec("c0"))
    pal = S.palette()
    wd = vt.World.dense(ctx, grid, pal)
    od = O.DenseWorld(grid, pal)
    rng = np.random.default_rng(3)
    n_hit = 0
    for k in range(40):
        for world, ow_, n in ((e["world"], ow, e["dim"]), (wd, od, 128)):
            eye = (rng.uniform(0, n), rng.uniform(0.6 * n, n), rng.uniform(-0.3 * n, 0))
            tgt = (rng.uniform(0.2 * n, 0.8 * n), rng.uniform(0, 0.4 * n), rng.uniform(0.2 * n, 0.8 * n))
            cam = vcam.look_at(eye, tgt, 70.0)
            g, o = world.pick(cam, 64), O.pick(ow_, cam, 64)
            for f in ("data", "face", "steps"):
                assert getattr(g, f) == getattr(o, f), (k, f)
            for f in ("hit_pos", "normal", "place", "remove"):
                assert list(getattr(g, f)) == list(getattr(o, f)), (k, f)
            n_hit += g.data != 0
    assert n_hit > 20
    wd.close()


@pytest.mark.parametrize("traversal", [vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8])
def test_dense_edits_render_like_edited_grid(ctx, traversal):
    import torch
    grid = S.generate(S.spec("c0"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    n0 = vt.build_compact_index(ctx, w)
    assert n0 == np.count_nonzero(grid)
    rng = np.random.default_rng(17 + traversal)
    exp = grid.copy()
    for batch in range(3):
        n = 20000
        xyz = rng.integers(0, 128, (n, 3)).astype(np.uint32)
        # carve a slab and drop blocks: clears of solid voxels and new solids
        xyz[: n // 2, 1] = rng.integers(30, 60, n // 2)
        val = np.where(rng.random(n) < 0.5, 0, rng.integers(1, 9, n)).astype(np.uint32)
        xyz[-50:] = xyz[:50]  # repeated positions: the later edit wins
        for (x, y, z), v in zip(xyz, val):
            exp[z, y, x] = v
        log(f"dense edits batch {batch}")
        w.set_voxels(xyz, val)
    assert not np.array_equal(exp, grid)
    sample = rng.integers(0, 128, (5000, 3)).astype(np.uint32)
    assert np.array_equal(w.get_voxels(sample), exp[sample[:, 2], sample[:, 1], sample[:, 0]])
    assert vt.build_compact_index(ctx, w) == np.count_nonzero(exp)
    cam, W, H = S.views("c0")[0]
    f = vt.Frame.alloc(W, H, gbuffer=True, steps=True)
    vt.render(ctx, w, cam, f, vt.params(W, H, 512, 512, S.SUN, 0, traversal))
    torch.cuda.synchronize()
    o = O.render(O.DenseWorld(exp, pal), cam, W, H, 512, 512, S.SUN, traversal=traversal)
    gh = f.hits.cpu().numpy().view(np.uint32)
    assert np.array_equal(gh, o["hits"]), np.sum(np.any(gh != o["hits"], -1))
    assert np.array_equal(f.steps.cpu().numpy().view(np.uint32), o["steps"])
    w.close()


def test_edit_argument_errors(ctx):
    grid = np.zeros((16, 16, 16), np.uint8)
    w = vt.World.dense(ctx, grid, S.palette())
    with pytest.raises(vt.VtError) as e:
        w.set_voxels([[16, 0, 0]], [1])
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    with pytest.raises(vt.VtError) as e:
        w.set_voxels([[1, 2, 3]], [256])
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    with pytest.raises(vt.VtError) as e:
        w.get_voxels([[0, 0, 99]])
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    assert w.get_voxels([[1, 2, 3]])[0] == 0  # nothing was written by the failed calls
    w.close()


def test_reads_in_staging_chunks(ctx):
    """vt_world_get_voxels stages through the context's buffer 16,384 voxels
    at a time: a read of 3 x that plus a ragged tail equals the grid, on
    both layouts, and a pick repeated 200 times gives one answer."""
    grid = S.generate(S.spec("c0"))
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    rng = np.random.default_rng(9)
    xyz = rng.integers(0, 128, (3 * 16384 + 123, 3)).astype(np.uint32)
    log("dense read of 49,275 voxels")
    assert np.array_equal(w.get_voxels(xyz), grid[xyz[:, 2], xyz[:, 1], xyz[:, 0]])
    import brickmap_util as BU
    dim, chunks, pool, atlas = BU.embedding_a(grid, pal)
    wb = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    ow = O.BrickmapWorld(dim, chunks, pool, atlas)
    got = wb.get_voxels(xyz)
    exp = np.array([0 if chunks[(x // 8) + (dim // 8) * ((y // 8) + (dim // 8) * (z // 8))] == 0
                    else pool[(chunks[(x // 8) + (dim // 8) * ((y // 8) + (dim // 8) * (z // 8))]
                               - 1) * 512 + (x % 8) + 8 * ((y % 8) + 8 * (z % 8))]
                    for x, y, z in xyz], np.uint32)
    assert np.array_equal(got, exp)
    cam = vcam.look_at((64.0, 110.0, -20.0), (64.0, 40.0, 64.0), 70.0)
    o = O.pick(ow, cam, 64)
    for _ in range(200):
        g = wb.pick(cam, 64)
        assert (g.data, g.face, g.steps) == (o.data, o.face, o.steps)
    wb.close()
    w.close()


def test_failed_brickmap_edit_leaves_world_unchanged(ctx, monkeypatch):
    """A pool growth refused by the VT_POOL_LIMIT_BYTES budget fails the
    call before anything changes: the chunk table and pool read back as
    before, and the next successful edit allocates the next block (no two
    chunks share a block)."""
    atlas = atlas_fixture()
    w = vt.World.empty_brickmap(ctx, 64, atlas)
    w.set_voxels([[1, 1, 1]], [3 | (1 << 28)])  # one chunk, one block
    c0, p0 = w.read_brickmap()
    assert p0.size == 512
    # 40 new chunks need 41 blocks: the pool would grow to 64 blocks (128 KiB)
    xyz = np.array([[8 * (i % 8), 8 * (i // 8), 0] for i in range(1, 41)], np.uint32)
    vals = np.full(len(xyz), 5 | (1 << 28), np.uint32)
    monkeypatch.setenv("VT_POOL_LIMIT_BYTES", str(16 * 2048))
    with pytest.raises(vt.VtError) as e:
        w.set_voxels(xyz, vals)
    assert e.value.status == vt._native.VT_ERR_OUT_OF_MEMORY
    c1, p1 = w.read_brickmap()
    assert np.array_equal(c0, c1) and np.array_equal(p0, p1)
    monkeypatch.delenv("VT_POOL_LIMIT_BYTES")
    w.set_voxels(xyz[:2], vals[:2])
    c2, p2 = w.read_brickmap()
    used = c2[c2 != 0]
    assert sorted(used.tolist()) == [1, 2, 3] and p2.size == 3 * 512
    assert np.array_equal(w.get_voxels(np.vstack([[[1, 1, 1]], xyz[:2]])),
                          np.array([3 | (1 << 28), 5 | (1 << 28), 5 | (1 << 28)], np.uint32))
    w.close()
