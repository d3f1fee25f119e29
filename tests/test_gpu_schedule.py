"""The tile schedule (vt_frame.schedule) through the C-ABI.

A render into a frame with a schedule buffer launches the frame's blocks in
the cost order the previous render into the same frame measured (the first
render, and renders into frames without a buffer, use the fixed order).
Which launch slot renders which image block is all that changes, so the bar
is bit-exact: every image of every render equals the fixed-order render of
the same view, render after render, and the record the buffer holds after a
render lists every block of the frame exactly once.
"""
import numpy as np
import pytest

import vt
from vt import _native as N
from vt import scenes as S

pytestmark = pytest.mark.gpu

HEAD, LINE, COUNTS, COPY, XCD_DONE = 512, 32, 4096, 4, 256
IMAGES = ("shaded", "hits", "color", "normal", "position", "illumination")


def tile_log2(ctx, W, H):
    """vt_render's tile shape: 16x4 when the frame has at least 8 blocks of
    8x8 tiles per CU, else 8x8."""
    cu = ctx.info()["compute_units"]
    return 4 if (-(-W // 16)) * (-(-H // 16)) >= 8 * cu else 3


def nblocks_of(W, H, lw):
    tw, th = 1 << lw, 64 >> lw
    return (-(-W // (2 * tw))) * (-(-H // (2 * th)))


def tag_of(W, H, lw):
    m = 0xFFFFFFFF
    return 0x80000000 | (((W * 0x9E3779B1) & m) ^ ((H * 0x85EBCA77) & m) ^ ((lw * 0xC2B2AE3D) & m))


def xcd_remap(i, n, strip=4):
    """The fixed order: launch slot i -> image block (strips of 4 blocks
    dealt to the 8 XCDs in turn; slots past the last full round map to
    themselves)."""
    rnd = 8 * strip
    if i >= (n // rnd) * rnd:
        return i
    x, j = i % 8, i // 8
    return ((j // strip) * 8 + x) * strip + j % strip


def check_record(f, W, H, lw, label, first=False):
    """The buffer after a render: the geometry tag of this frame, reset
    counters, and the record of the render in slot [0]: per XCD, class
    counts (copied into the head) summing to the XCD's block count, and
    lists that hold exactly the image blocks the fixed order gives that XCD.
    first: the render was the buffer's first at this size, which clears the
    slots and records nothing (all counts 0)."""
    s = f.schedule.cpu().numpy().view(np.uint32)
    nb = nblocks_of(W, H, lw)
    C = -(-N.schedule_blocks(W, H) // 8)
    assert s[1] == tag_of(W, H, lw), f"{label}: tag {s[1]:#x}"
    assert s[2] == 0 and all(s[XCD_DONE + LINE * x] == 0 for x in range(8)), f"{label}: counters"
    r = int(s[0])
    assert r < 3, label
    slot = HEAD + r * (COUNTS + 128 * C)
    counts = np.array([int(s[slot + q * LINE]) for q in range(128)]).reshape(8, 16)
    assert counts.ravel().tolist() == [int(v) for v in s[COPY:COPY + 128]], \
        f"{label}: head copy of the counts"
    if first:
        assert counts.sum() == 0, f"{label}: a first render recorded {counts.sum()} blocks"
        return counts.sum(0)
    for x in range(8):
        want = sorted(xcd_remap(i, nb) for i in range(x, nb, 8))
        assert counts[x].sum() == len(want), f"{label}: XCD {x} recorded {counts[x].sum()}"
        got = np.concatenate([s[slot + COUNTS + (x * 16 + k) * C:
                                slot + COUNTS + (x * 16 + k) * C + counts[x, k]]
                              for k in range(16)])
        assert sorted(got.tolist()) == want, f"{label}: XCD {x} lists other blocks"
    return counts.sum(0)


def snapshot(f):
    return {k: getattr(f, k).clone() for k in IMAGES + ("steps", "stats")
            if getattr(f, k) is not None}


def assert_same(a, b, label):
    import torch
    for k in a:
        assert torch.equal(a[k], b[k]), f"{label}: {k} differs from the fixed order"


def poison(f):
    from test_gpu_parity import poison as p
    return p(f)


@pytest.fixture(scope="module")
def c0_world(ctx):
    grid = S.generate(S.spec("c0"))
    w = vt.World.dense(ctx, grid, S.palette())
    yield grid, w
    w.close()


def reference(ctx, world, cam, W, H, prm, **kw):
    import torch
    f = poison(vt.Frame.alloc(W, H, schedule=False, **kw))
    if f.stats is not None:
        f.stats.zero_()
    vt.render(ctx, world, cam, f, prm)
    torch.cuda.synchronize()
    return snapshot(f)


def renders_match(ctx, world, cam, W, H, prm, n=6, label="", **kw):
    """n renders into one frame with a schedule, each against the fixed
    order; the record after each is checked.  The schedule is taken for a
    step cap of at least 512; below, the buffer is left as it was (zero)."""
    import torch
    ref = reference(ctx, world, cam, W, H, prm, **kw)
    f = vt.Frame.alloc(W, H, **kw)
    lw = tile_log2(ctx, W, H)
    taken = max(prm.max_steps_primary, prm.max_steps_shadow) >= 512
    counts = None
    for i in range(n):
        poison(f)
        if f.stats is not None:
            f.stats.zero_()
        vt.render(ctx, world, cam, f, prm)
        torch.cuda.synchronize()
        assert_same(snapshot(f), ref, f"{label} render {i}")
        if taken:
            counts = check_record(f, W, H, lw, f"{label} render {i}", first=i == 0)
        else:
            assert int(f.schedule.abs().sum().item()) == 0, f"{label}: short caps used the buffer"
    return f, counts


@pytest.mark.parametrize("W,H", [(1280, 720), (256, 256), (333, 517)])
def test_dense_renders_equal_fixed_order(ctx, c0_world, W, H):
    """C0 world at three sizes: 16x4 tiles with more blocks than the GPU
    holds at once, 8x8 tiles, and a size that cuts blocks at both edges;
    G-buffer, step counts and statistics included."""
    _, w = c0_world
    cam, _, _ = S.views("c0")[0]
    prm = vt.params(W, H, 512, 512, S.SUN)
    f, counts = renders_match(ctx, w, cam, W, H, prm, label=f"dense {W}x{H}", gbuffer=True,
                              steps=True, stats=True)
    # the free-region test on (step counts not stored), Lambert and present
    for flags, kw in ((vt.VT_FLAG_FREE_MAP_ON, {}), (vt.VT_FLAG_PRESENT, {"display": True})):
        prm2 = vt.params(W, H, 512, 512, S.SUN, flags)
        renders_match(ctx, w, cam, W, H, prm2, n=4, label=f"dense {W}x{H} flags {flags}", **kw)
    print(f"[schedule] {W}x{H}: class counts {counts}", flush=True)


def test_dense_lambert_equal_fixed_order(ctx, c0_world):
    import torch
    grid, w = c0_world
    cam, _, _ = S.views("c0")[0]
    W, H = 1280, 720
    albedo = torch.rand(grid.size * 3, generator=torch.Generator().manual_seed(5)).cuda()
    prm = vt.params(W, H, 512, 512, S.SUN, vt.VT_FLAG_LAMBERT, 0, 0.2, albedo)
    renders_match(ctx, w, cam, W, H, prm, n=4, label="lambert")


@pytest.mark.parametrize("traversal", [vt.VT_TRAVERSAL_BITMASK, vt.VT_TRAVERSAL_BITMASK8])
def test_bitmask_traversals_equal_fixed_order(ctx, c0_world, traversal):
    _, w = c0_world
    cam, _, _ = S.views("c0")[0]
    W, H = 1280, 720
    prm = vt.params(W, H, 512, 512, S.SUN, 0, traversal)
    renders_match(ctx, w, cam, W, H, prm, n=3, label=f"traversal {traversal}")


def test_reference_layout_equal_fixed_order(ctx, c0_world):
    import brickmap_util as BU
    grid, _ = c0_world
    dim, chunks, pool, atlas = BU.embedding_a(grid, S.palette())
    wb = vt.World.brickmap(ctx, dim, chunks, pool, atlas)
    cam, _, _ = S.views("c0")[0]
    W, H = 1280, 720
    for caps, flags in (((512, 512), 0), ((192, 48), vt.VT_FLAG_GBUF_UNORM8)):
        prm = vt.params(W, H, caps[0], caps[1], S.SUN, flags)
        renders_match(ctx, wb, cam, W, H, prm, n=4, label=f"reference layout {caps}",
                      gbuffer=True)
    wb.close()


def test_cameras_alternating_in_one_frame(ctx, c0_world):
    """Two views rendered in turn into one frame: each render launches in
    the order the other view measured, and still equals its fixed order."""
    import torch
    _, w = c0_world
    W, H = 1280, 720
    cams = [S.views("c0")[0][0], vt.look_at((10.0, 90.0, 120.0), (100.0, 40.0, 20.0), 80.0)]
    prm = vt.params(W, H, 512, 512, S.SUN)
    refs = [reference(ctx, w, c, W, H, prm) for c in cams]
    f = vt.Frame.alloc(W, H)
    for i in range(6):
        poison(f)
        vt.render(ctx, w, cams[i % 2], f, prm)
        torch.cuda.synchronize()
        assert_same(snapshot(f), refs[i % 2], f"alternating render {i}")
        check_record(f, W, H, tile_log2(ctx, W, H), f"alternating render {i}", first=i == 0)


def test_buffer_used_at_another_size(ctx, c0_world):
    """A buffer that last served another image size is restarted (fixed
    order, slots cleared) and then orders renders of the new size."""
    import torch
    _, w = c0_world
    cam, _, _ = S.views("c0")[0]
    big = vt.Frame.alloc(1280, 720)
    sched = big.schedule
    for (W, H) in ((1280, 720), (1280, 720), (640, 360), (640, 360), (640, 360), (1280, 720),
                   (1280, 720)):
        prm = vt.params(W, H, 512, 512, S.SUN)
        ref = reference(ctx, w, cam, W, H, prm)
        f = vt.Frame.alloc(W, H, schedule=False)
        f.schedule = sched
        poison(f)
        vt.render(ctx, w, cam, f, prm)
        torch.cuda.synchronize()
        assert_same(snapshot(f), ref, f"{W}x{H}")
        s = sched.cpu().numpy().view(np.uint32)
        assert s[1] == tag_of(W, H, tile_log2(ctx, W, H))


def test_graph_replays_rotate_the_record(ctx, c0_world):
    """vt_render captured in a graph: the record's slot lives in the buffer,
    so every replay reads the last replay's record and writes the next; each
    replay equals the fixed order (eager work and syncs between replays)."""
    import torch
    _, w = c0_world
    cam, _, _ = S.views("c0")[0]
    W, H = 1280, 720
    prm = vt.params(W, H, 512, 512, S.SUN)
    ref = reference(ctx, w, cam, W, H, prm)
    f = vt.Frame.alloc(W, H)
    vt.render(ctx, w, cam, f, prm)  # eager first render: the buffer starts
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            vt.render(ctx, w, cam, f, prm)
    torch.cuda.synchronize()
    slots = []
    for i in range(5):
        poison(f)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert_same(snapshot(f), ref, f"replay {i}")
        check_record(f, W, H, tile_log2(ctx, W, H), f"replay {i}")
        slots.append(int(f.schedule[0].item()))
        _ = (f.hits.float().sum()).item()  # eager work between replays
    assert len(set(slots)) == 3, f"the record's slot did not rotate: {slots}"
