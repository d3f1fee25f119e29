"""Edge cases of the render path on the GPU vs the oracle, through the C-ABI.

Worlds: empty, completely solid, a checkerboard, single voxels in corners,
random sparse grids, the smallest dense size (16^3); cameras: outside looking
away (all sky), exactly on lattice points, inside a solid voxel, looking
exactly along each axis (zero direction components: the 0.001
substitution), grazing a face; images 1x1, 1xN, Nx1 and ragged tile edges;
step caps 0, 1, 2 and a zero shadow cap; both traversals and the reference
layout of the same worlds; frames that fill every CU's block slots.  The bar is the one of the main parity tests:
hit records, step counts, G-buffer position/normal/illumination and the
statistics counters bit-exact, colours within 1e-5.
"""
import numpy as np
import pytest

import oracle_ffi as O
import vt
from test_gpu_parity import compare, gpu_render
from vt import brickmap as BM
from vt import scenes as S

pytestmark = pytest.mark.gpu


def worlds():
    rng = np.random.default_rng(21)
    out = {}
    out["empty16"] = np.zeros((16, 16, 16), np.uint8)
    out["solid16"] = np.full((16, 16, 16), 3, np.uint8)
    z, y, x = np.indices((32, 32, 32))
    out["checker32"] = (((x + y + z) & 1) * (1 + (x % 8))).astype(np.uint8)
    c = np.zeros((32, 32, 32), np.uint8)
    for (a, b, d) in [(0, 0, 0), (31, 0, 0), (0, 31, 0), (0, 0, 31), (31, 31, 31), (15, 16, 15)]:
        c[d, b, a] = 1 + (a + b + d) % 8
    out["corners32"] = c
    g = (rng.random((64, 64, 64)) < 0.02) * rng.integers(1, 9, (64, 64, 64))
    out["sparse64"] = g.astype(np.uint8)
    return out


WORLDS = worlds()


def cameras(n):
    c = n / 2.0
    return [
        ("outside_away", vt.look_at((c, c, -3 * n), (c, c, -6 * n))),
        ("lattice_point", vt.look_at((c, c, c), (c + 3.0, c - 2.0, c + 5.0))),
        ("inside_voxel", vt.look_at((c + 0.3, c + 0.6, c + 0.45), (0.0, 0.0, 0.0))),
        ("axis_pos_x", vt.look_at((-4.0, c + 0.5, c + 0.5), (n + 4.0, c + 0.5, c + 0.5))),
        ("axis_neg_y", vt.look_at((c + 0.25, n + 6.0, c + 0.75), (c + 0.25, -6.0, c + 0.75),
                                  up=(0.0, 0.0, 1.0))),
        ("axis_neg_z", vt.look_at((c + 0.5, c + 0.5, n + 5.0), (c + 0.5, c + 0.5, -5.0))),
        ("grazing_face", vt.look_at((-2.0, float(n // 2), -2.0), (n + 2.0, float(n // 2), n + 2.0))),
        ("corner_view", vt.look_at((-1.0, -1.0, -1.0), (float(n), float(n), float(n)))),
    ]


SIZES = [(1, 1), (1, 7), (13, 1), (65, 3), (40, 33)]


@pytest.mark.parametrize("wname", list(WORLDS))
@pytest.mark.parametrize("traversal", [vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8])
def test_edge_worlds_cameras_sizes(ctx, wname, traversal):
    grid = WORLDS[wname]
    pal = S.palette()
    n = grid.shape[0]
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    for cname, cam in cameras(n):
        for (W, H) in SIZES:
            label = f"{wname} {cname} {W}x{H} trav={traversal}"
            g = gpu_render(ctx, w, cam, W, H, 4 * n, 4 * n, traversal=traversal)
            o = O.render(ow, cam, W, H, 4 * n, 4 * n, S.SUN, traversal=traversal)
            compare(g, o, label=label)
    w.close()


@pytest.mark.parametrize("caps", [(0, 0), (1, 0), (2, 1), (5, 0), (6, 6), (10, 11), (64, 2)])
def test_edge_step_caps(ctx, caps):
    grid = WORLDS["sparse64"]
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    for cname, cam in cameras(64)[1:4]:
        for traversal in (vt.VT_TRAVERSAL_FLAT, vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8):
            g = gpu_render(ctx, w, cam, 40, 33, caps[0], caps[1], traversal=traversal)
            o = O.render(ow, cam, 40, 33, caps[0], caps[1], S.SUN, traversal=traversal)
            compare(g, o, label=f"caps {caps} {cname} trav={traversal}")
    w.close()


@pytest.mark.parametrize("wname", ["solid16", "checker32", "corners32", "sparse64"])
def test_edge_worlds_reference_layout(ctx, wname):
    grid = WORLDS[wname]
    pal = S.palette()
    n = grid.shape[0]
    dim, chunks, pool, atlas = BM.from_dense(grid, pal)
    if pool.size == 0:
        pytest.skip("empty world has no pool block")
    w = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    ow = O.BrickmapWorld(dim, chunks, pool, atlas)
    for cname, cam in cameras(n):
        g = gpu_render(ctx, w, cam, 40, 33, 8 * n, 8 * n)
        o = O.render(ow, cam, 40, 33, 8 * n, 8 * n, S.SUN)
        compare(g, o, stats=False, label=f"brickmap {wname} {cname}")
    w.close()


def test_edge_empty_reference_layout(ctx):
    """An empty map (no pool block at all): every ray is sky."""
    pal = S.palette()
    dim, chunks, pool, atlas = BM.from_dense(WORLDS["empty16"], pal)
    w = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    ow = O.BrickmapWorld(dim, chunks, np.zeros(512, np.uint32), atlas)
    cam = cameras(16)[1][1]
    g = gpu_render(ctx, w, cam, 40, 33, 128, 128)
    o = O.render(ow, cam, 40, 33, 128, 128, S.SUN)
    compare(g, o, stats=False, label="empty reference layout")
    assert np.all(g["hits"][..., 0] == 0xFFFFFFFF)
    w.close()


# Suns and cameras on exact diagonals: crossing times tie on two or three
# axes at once, at every step, for rays that start on lattice points (all
# shadow rays do).  These drive the tie rule (z > y > x), the within >= 8
# correction on one or two axes at a time, and the empty-space loops' exact
# replays (for a sun with no negative component, two axes at t <= 0 stop an
# iteration).
TIE_SUNS = [(0.5, 0.5, 0.5), (1.0, 1.0, 0.0), (0.6, 0.8, 0.0), (0.25, 0.5, 0.25),
            (-0.5, 0.5, 0.5), (-1.0, -1.0, -1.0)]


@pytest.mark.parametrize("sun", TIE_SUNS)
@pytest.mark.parametrize("wname", ["checker32", "sparse64"])
def test_edge_tied_suns(ctx, wname, sun):
    grid = WORLDS[wname]
    pal = S.palette()
    n = grid.shape[0]
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    for cname, cam in cameras(n)[1:3] + cameras(n)[6:8]:
        g = gpu_render(ctx, w, cam, 40, 33, 8 * n, 8 * n, sun=sun)
        o = O.render(ow, cam, 40, 33, 8 * n, 8 * n, sun)
        compare(g, o, label=f"{wname} {cname} sun={sun}")
    w.close()


def test_edge_diagonal_cameras(ctx):
    """Primary rays along exact diagonals from lattice points."""
    grid = WORLDS["sparse64"]
    pal = S.palette()
    w = vt.World.dense(ctx, grid, pal)
    ow = O.DenseWorld(grid, pal)
    cams = [
        ("diag_ppp", vt.look_at((1.0, 1.0, 1.0), (63.0, 63.0, 63.0))),
        ("diag_npn", vt.look_at((63.0, 1.0, 63.0), (1.0, 63.0, 1.0))),
        ("diag_xy", vt.look_at((0.0, 0.0, 32.0), (64.0, 64.0, 32.0), up=(0.0, 0.0, 1.0))),
        ("diag_mid", vt.look_at((32.0, 32.0, 32.0), (40.0, 40.0, 40.0))),
    ]
    for cname, cam in cams:
        for (W, H) in [(1, 1), (41, 41)]:
            g = gpu_render(ctx, w, cam, W, H, 512, 512, sun=(0.5, 0.5, 0.5))
            o = O.render(ow, cam, W, H, 512, 512, (0.5, 0.5, 0.5))
            compare(g, o, label=f"{cname} {W}x{H}")
    w.close()


def test_edge_large_frames(ctx):
    """Frames with enough blocks (16x16 pixels each) to fill every CU's 8
    block slots, shadow caps around the empty-space iteration's length and
    long ones: tied suns (two- and three-axis ties every step), suns of
    either sign, lattice-point cameras, on the dense and the reference
    layout, against the oracle."""
    W, H = 1280, 720
    cus = ctx.info()["compute_units"]
    assert ((W + 15) // 16) * ((H + 15) // 16) >= 8 * cus
    pal = S.palette()
    cases = [("checker32", (0.5, 0.5, 0.5), 256), ("checker32", (-0.5, 0.5, 0.5), 256),
             ("sparse64", (1.0, 1.0, 0.0), 512), ("sparse64", S.SUN, 5),
             ("sparse64", S.SUN, 6), ("sparse64", (0.25, 0.5, 0.25), 7),
             ("sparse64", (-0.5, 0.7, -0.2), 11), ("corners32", S.SUN, 256)]
    for wname, sun, cap in cases:
        grid = WORLDS[wname]
        n = grid.shape[0]
        w = vt.World.dense(ctx, grid, pal)
        ow = O.DenseWorld(grid, pal)
        dim, chunks, pool, atlas = BM.from_dense(grid, pal)
        wb = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
        owb = O.BrickmapWorld(dim, chunks, pool, atlas)
        for cname, cam in (cameras(n)[1], cameras(n)[7]):
            g = gpu_render(ctx, w, cam, W, H, 8 * n, cap, sun=sun)
            o = O.render(ow, cam, W, H, 8 * n, cap, sun)
            compare(g, o, label=f"large {wname} {cname} sun={sun} shadow cap {cap}")
            # the reference layout: sub-voxel steps inside the hit block first
            g = gpu_render(ctx, wb, cam, W, H, 64 * n, 8 * cap, sun=sun)
            o = O.render(owb, cam, W, H, 64 * n, 8 * cap, sun)
            compare(g, o, stats=False, label=f"large brickmap {wname} {cname} sun={sun}")
        w.close()
        wb.close()


def test_edge_large_present_fused(ctx):
    """The fused present pass on a frame that fills the GPU (1280x720,
    shadow cap 512):
    shaded and display images equal to a plain render followed by
    vt_present, on the dense and the reference layout."""
    import torch
    W, H = 1280, 720
    grid = WORLDS["sparse64"]
    pal = S.palette()
    dim, chunks, pool, atlas = BM.from_dense(grid, pal)
    worlds = [vt.World.dense(ctx, grid, pal), vt.World.brickmap(ctx, dim, chunks, pool, atlas)]
    cam = cameras(64)[1][1]
    for w, cap in zip(worlds, (512, 4096)):
        f = vt.Frame.alloc(W, H, display=True)
        vt.render(ctx, w, cam, f, vt.params(W, H, cap, cap, S.SUN, vt.VT_FLAG_PRESENT))
        f2 = vt.Frame.alloc(W, H)
        vt.render(ctx, w, cam, f2, vt.params(W, H, cap, cap, S.SUN))
        out = torch.empty_like(f2.shaded)
        disp = torch.empty((H, W), dtype=torch.int32, device="cuda")
        vt.present(ctx, f2.shaded, out, disp)
        torch.cuda.synchronize()
        assert torch.equal(f.hits, f2.hits)
        assert torch.equal(out, f.shaded) and torch.equal(disp, f.display)
        w.close()
