"""How far traceMap's probes stray from the ray's line (CPU, the oracle).

With the free-region test on, vt_render ends a primary ray of a dense world
as a miss when the line e + d t (t >= 0, e the traversal's start, d the
direction after the zero substitution) misses the box of the nonempty
voxels widened by 2.5 + dim / 1000 voxels.  That is exact when every block
traceMap probes lies within the widening of the line: the walk drifts from
the line by 0.001 voxel per block boundary it crosses on a negative axis
(the 0.999 re-anchor, map.glsl:161), at most dim such crossings per axis
inside the grid, plus binary32 rounding, and a probe is the position's block
or the next one on a positive axis whose within reached 8.

Here the oracle's traceMap walks empty grids (every ray runs to the grid's
edge, the longest walks there are) and records every probed block; for each
the test computes the smallest widening of the block that the line meets and
checks that it stays far below the margin -- and that the drift the bound is
made for is really there (corner-to-corner rays with mixed signs stray by a
good part of dim / 1000 voxels).
"""
import numpy as np
import pytest

import oracle_ffi as O


def wide_margin(dim):
    return 2.5 + dim / 1000.0


def needed_widening(e, d, blocks, t_max):
    """Smallest w >= 0 such that the line e + d t, 0 <= t <= t_max, meets
    [b - w, b + 1 + w]^3, for each row b of blocks (float64, ternary search
    over t of the convex distance max_i dist(e_i + d_i t, [b_i, b_i + 1]))."""
    b = blocks.astype(np.float64)
    lo = np.zeros(len(b))
    hi = np.full(len(b), float(t_max))

    def f(t):
        p = e[None, :] + d[None, :] * t[:, None]
        return np.max(np.maximum(np.maximum(b - p, p - (b + 1.0)), 0.0), axis=1)

    for _ in range(80):
        m1 = lo + (hi - lo) / 3.0
        m2 = hi - (hi - lo) / 3.0
        left = f(m1) <= f(m2)
        hi = np.where(left, m2, hi)
        lo = np.where(left, lo, m1)
    return f(0.5 * (lo + hi))


def rays(dim, rng, count):
    """Starts inside the grid and directions of every octant: random ones,
    ones with zero components (the 0.001 substitution), shallow ones, and
    the worst case for the drift: corner to corner across the whole grid
    with mixed signs (on all-negative diagonals the drift runs along the
    line itself)."""
    out = []
    for k in range(count):
        kind = k % 4
        if kind == 0:
            e = rng.uniform(0.0, dim, 3)
            d = rng.normal(size=3)
        elif kind == 1:
            e = rng.uniform(0.0, dim, 3)
            d = rng.normal(size=3)
            d[rng.integers(0, 3)] = 0.0
        elif kind == 2:
            e = rng.uniform(0.0, dim, 3)
            d = rng.normal(size=3)
            d[rng.integers(0, 3)] *= 1e-3
        else:
            # across the grid with mixed signs: the re-anchor drift on the
            # negative axes points off the line
            sg = np.array([-1.0, 1.0, 1.0 if k % 8 == 3 else -1.0])[rng.permutation(3)]
            e = np.where(sg < 0, dim - rng.uniform(0.001, 0.5, 3), rng.uniform(0.001, 0.5, 3))
            d = sg * (1.0 + 0.2 * rng.uniform(-1.0, 1.0, 3))
        d = d / np.linalg.norm(d)
        out.append((e.astype(np.float32), d.astype(np.float32)))
    return out


@pytest.mark.parametrize("dim,count", [(16, 400), (64, 300), (256, 160), (1024, 40)])
def test_probes_stay_within_the_margin(dim, count):
    grid = np.zeros((dim, dim, dim), np.uint8)
    world = O.DenseWorld(grid, np.zeros(256, np.uint32))
    rng = np.random.default_rng(dim)
    worst = 0.0
    corner_worst = 0.0
    for k, (e, d) in enumerate(rays(dim, rng, count)):
        cap = 4 * dim
        h, path = O.trace(world, e, d, cap, path_cap=cap)
        assert h.steps < cap, "a walk of an empty grid ends at the grid's edge"
        blocks = path >> 3  # probed sub-voxel positions -> blocks
        d_eff = np.where(d == 0.0, np.float32(0.001), d).astype(np.float64)
        t_max = 4.0 * dim / np.min(np.abs(d_eff))
        w = needed_widening(e.astype(np.float64), d_eff, blocks, min(t_max, 1e7))
        worst = max(worst, float(w.max()))
        if k % 4 == 3:
            corner_worst = max(corner_worst, float(w.max()))
    print(f"[reach] dim {dim}: largest widening a probe needed {worst:.4f} voxel "
          f"(corner rays {corner_worst:.4f}), margin {wide_margin(dim):.3f}")
    # the argument's bound (drift dim / 1000 plus one block for the within
    # >= 8 probe) holds with room, and the margin is above it
    assert worst <= dim / 1000.0 + 1.0 + 1e-3
    assert worst + 1.0 < wide_margin(dim)
    # the re-anchor drift the bound accounts for is real: corner rays cross
    # every block boundary of their negative axes
    if dim >= 64:
        assert corner_worst > 0.25 * dim / 1000.0


def test_needed_widening_closed_forms():
    """The distance helper on lines whose distance to a block is known."""
    e = np.array([0.5, 3.25, 0.5])
    d = np.array([1.0, 0.0, 0.0])  # parallel to x, 2.25 above block y 0
    b = np.array([[5, 0, 0], [5, 3, 0], [5, 5, 0], [-3, 0, 0]])
    w = needed_widening(e, d, b, 100.0)
    # the block behind the start (x -3) is reached only at t = -2.5 (t >= 0):
    # the nearest point is the start, 2.5 voxels past x -2 and 2.25 above
    assert np.allclose(w, [2.25, 0.0, 1.75, 2.5])
    d2 = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)  # diagonal through (0.5,0.5)
    w2 = needed_widening(np.array([0.5, 0.5, 0.5]), d2, np.array([[3, 0, 0], [10, 10, 0]]),
                         100.0)
    # block (3, 0): the line x - y = 0 against [3, 4] x [0, 1]: distance
    # min over t of max(3 - x, y - 1) with x = y -> 1 at x = y = 2
    assert np.allclose(w2, [1.0, 0.0])


def blind_steps_f32(e, d, lo, hi):
    """The number of leading steps whose lookups the traversal skips (the
    kernel's blind_steps, binary32): the block boundaries the line crosses
    before it enters the reach box [lo, hi], less 13; None when the line
    misses the reach box (the ray ends before its walk)."""
    e = e.astype(np.float32)
    d = np.where(d == 0.0, np.float32(0.001), d).astype(np.float32)
    inv = (np.float32(1.0) / d).astype(np.float32)
    a = ((lo.astype(np.float32) - e) * inv).astype(np.float32)
    b = ((hi.astype(np.float32) - e) * inv).astype(np.float32)
    tn = np.float32(max(0.0, float(np.max(np.minimum(a, b)))))
    if tn > np.float32(min(3.0e38, float(np.min(np.maximum(a, b))))):
        return None
    p = (e + d * tn).astype(np.float32)
    return int(np.sum(np.abs(np.floor(p).astype(np.int64) - np.floor(e).astype(np.int64)))) - 13


@pytest.mark.parametrize("dim", [64, 256, 1024])
def test_blind_steps_probe_outside_the_box(dim):
    """The steps a primary ray takes before its lookups are made (toward the
    box of the nonempty voxels from outside its reach) probe only blocks
    outside that box: the oracle's walk, probe by probe, against boxes of
    every size, with rays aimed at the box, grazing it and crossing it at
    every angle.  Also reports how many probes the bound leaves to the
    lookups before the first one inside the box."""
    grid = np.zeros((dim, dim, dim), np.uint8)
    world = O.DenseWorld(grid, np.zeros(256, np.uint32))
    rng = np.random.default_rng(3 * dim)
    m = wide_margin(dim)
    slack_min, checked = 10 ** 9, 0
    for k in range(300):
        size = rng.integers(1, dim // 2, 3)
        blo = rng.integers(0, dim - size + 1)
        bhi = blo + size - 1  # inclusive block box
        lo, hi = blo - m, bhi + 1.0 + m
        e = rng.uniform(0.0, dim, 3)
        if np.all((e >= lo) & (e <= hi)):
            continue
        tgt = rng.uniform(blo - 3.0, bhi + 4.0)
        d = tgt - e
        if k % 5 == 0:
            d[rng.integers(0, 3)] = 0.0
        d = (d / np.linalg.norm(d)).astype(np.float32)
        e = e.astype(np.float32)
        nb = blind_steps_f32(e, d, lo, hi)
        if nb is None or nb <= 0:
            continue
        cap = 4 * dim
        h, path = O.trace(world, e, d, cap, path_cap=cap)
        blocks = path >> 3
        inside = np.all((blocks >= blo) & (blocks <= bhi), axis=1)
        first = int(np.argmax(inside)) if inside.any() else len(blocks)
        assert first > nb, (f"dim {dim} ray {k}: probe {first} is inside the box but the "
                            f"first {nb} steps skip their lookups")
        if inside.any():
            slack_min = min(slack_min, first - nb)
        checked += 1
    print(f"[blind] dim {dim}: {checked} rays; fewest probes between the skipped steps and "
          f"the first probe inside the box: {slack_min}")
    assert checked > 100
