"""The reference world generator on the GPU (vt_world_procgen), through the C-ABI.

* The chunk table and block pool equal, word for word, those the oracle's
  literal restatement of procgen.zig leaves in a VoxelBrickmap (allocation
  order included): the engine's 512^3 map at the offsets Game.init uses
  (0, 0), at offsets that reach negative noise coordinates, at 1024^3, at
  504^3 (a map that is not a power of two: no occupancy bytes), and at
  small maps (no trees).  A map on which a tree would leave the bounds
  (320^3) is rejected and left unchanged, as the oracle flags it.
* The generated world (with the engine's 29-slot model atlas) renders
  exactly like a world created from the oracle's buffers, and like the
  oracle's render on a band of rows: the occupancy bytes of the fast
  traversal were rebuilt.  The engine's own first frame (Game.init's camera)
  and a turned view, 1920x1080, equal the oracle's full frames.
* Argument errors: dense worlds, non-empty worlds, NaN offsets.
"""
import numpy as np
import pytest

import oracle_ffi as O
import vt
from test_gpu_edit import atlas_fixture
from vt import camera as vcam
from vt import scenes as S

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[procgen] {msg}", flush=True)


@pytest.fixture(scope="module")
def atlas():
    return atlas_fixture()


@pytest.mark.parametrize("dim,ox,oy", [(512, 0.0, 0.0), (512, -37.25, 1000.5), (1024, 0.0, 0.0),
                                       (504, 0.0, 0.0), (256, 40.0, -20.0), (64, 3.0, 7.0),
                                       (16, 0.0, 0.0)])
def test_procgen_equals_oracle(ctx, atlas, dim, ox, oy):
    rc, chunks, pool = O.procgen(dim, ox, oy)
    assert rc == 0
    log(f"dim {dim} offsets ({ox}, {oy}): {pool.size // 512} blocks")
    w = vt.World.empty_brickmap(ctx, dim, atlas)
    w.procgen(ox, oy)
    gc, gp = w.read_brickmap()
    assert np.array_equal(gc, chunks), f"chunk table: {np.count_nonzero(gc != chunks)} differ"
    assert gp.size == pool.size
    bad = np.flatnonzero(gp != pool)
    assert bad.size == 0, f"pool: {bad.size} words differ, first at {bad[:5]}"
    w.close()


def test_procgen_outside_map_rejected(ctx, atlas):
    rc, _, _ = O.procgen(320)
    assert rc == 2  # the generator plants a tree whose leaves leave a 320^3 map
    w = vt.World.empty_brickmap(ctx, 320, atlas)
    with pytest.raises(vt.VtError) as e:
        w.procgen()
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    gc, gp = w.read_brickmap()
    assert gp.size == 0 and not gc.any()
    w.close()


def test_procgen_world_renders(ctx, atlas):
    import torch
    rc, chunks, pool = O.procgen(512)
    assert rc == 0
    w = vt.World.empty_brickmap(ctx, 512, atlas)
    w.procgen()
    ref = vt.World.brickmap(ctx, 512, chunks, pool, atlas)
    ow = O.BrickmapWorld(512, chunks, pool, atlas)
    # Game.init's player stands at (256, 22, 256), the camera 3 blocks up
    views = [((256.0, 25.0, 256.0), (300.0, 18.0, 290.0)),
             ((-40.0, 90.0, -40.0), (256.0, 20.0, 256.0)),
             ((480.0, 60.0, 100.0), (200.0, 10.0, 300.0))]
    W, H = 320, 180
    for k, (eye, tgt) in enumerate(views):
        cam = vcam.look_at(eye, tgt, 70.0)
        prm = vt.params(W, H, 2048, 512, S.SUN)
        f = vt.Frame.alloc(W, H, steps=True)
        g = vt.Frame.alloc(W, H, steps=True)
        log(f"view {k}")
        vt.render(ctx, w, cam, f, prm)
        vt.render(ctx, ref, cam, g, prm)
        torch.cuda.synchronize()
        fh = f.hits.cpu().numpy().view(np.uint32)
        assert np.array_equal(fh, g.hits.cpu().numpy().view(np.uint32)), f"view {k}"
        assert np.array_equal(f.steps.cpu().numpy(), g.steps.cpu().numpy())
        # rays ended early (free-region test, box of the generated blocks):
        # the same hits
        e = vt.Frame.alloc(W, H)
        vt.render(ctx, w, cam, e, vt.params(W, H, 2048, 512, S.SUN, vt.VT_FLAG_FREE_MAP_ON))
        torch.cuda.synchronize()
        assert np.array_equal(fh, e.hits.cpu().numpy().view(np.uint32)), f"view {k} early end"
        rows = (80, 96)
        o = O.render(ow, cam, W, H, 2048, 512, S.SUN, rows=rows, gbuffer=False)
        assert np.array_equal(fh[rows[0]:rows[1]], o["hits"][rows[0]:rows[1]]), f"view {k}"
        assert (fh[..., 0] != 0xFFFFFFFF).mean() > 0.3
    w.close()
    ref.close()


def test_procgen_engine_first_frames(ctx, atlas):
    """The engine's own first frame: Game.init's world and camera (player at
    (256, 22, 256), camera 3 blocks up, identity rotation, fov pi/2,
    camera.zig:5-10, game.zig:40,212), then a turned view; full 1920x1080
    frames against the oracle."""
    import math

    import torch
    from vt import _native as N
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
    W, H = 1920, 1080
    for k, cam in enumerate((start, turned)):
        prm = vt.params(W, H, 2048, 2048, S.SUN)
        f = vt.Frame.alloc(W, H, steps=True)
        log(f"engine frame {k}")
        vt.render(ctx, w, cam, f, prm)
        torch.cuda.synchronize()
        o = O.render(ow, cam, W, H, 2048, 2048, S.SUN, gbuffer=False)
        fh = f.hits.cpu().numpy().view(np.uint32)
        bad = np.any(fh != o["hits"], -1)
        assert not bad.any(), f"frame {k}: {bad.sum()} hit records differ"
        assert np.array_equal(f.steps.cpu().numpy().view(np.uint32), o["steps"])
        assert np.max(np.abs(f.shaded.cpu().numpy() - o["shaded"])) <= 1e-5
        assert (fh[..., 0] != 0xFFFFFFFF).mean() > 0.2
    w.close()


def test_procgen_argument_errors(ctx, atlas):
    d = vt.World.dense(ctx, np.zeros((16, 16, 16), np.uint8), S.palette())
    with pytest.raises(vt.VtError) as e:
        d.procgen()
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    d.close()
    w = vt.World.empty_brickmap(ctx, 64, atlas)
    with pytest.raises(vt.VtError) as e:
        w.procgen(float("nan"), 0.0)
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    w.set_voxels([[1, 2, 3]], [5 | (1 << 28)])
    with pytest.raises(vt.VtError) as e:
        w.procgen()
    assert e.value.status == vt._native.VT_ERR_INVALID_ARGUMENT
    assert w.get_voxels([[1, 2, 3]])[0] == 5 | (1 << 28)
    w.close()
