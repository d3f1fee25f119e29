"""Pins the CPU restatement (oracle/) to what the mathematics fixes.

The reference path is GLSL with no tests, fixtures or golden images of its
own, and no GL context can be created in this image, so every oracle function
is pinned here by checks that do not retype its formula:

  raygen (primary.comp.glsl:31-39)   float64 closed form; F4 4-component
                                      normalize (centre pixel norm 1/sqrt 2);
                                      non-square images pin the aspect term.
  intersectAABB (map.glsl:21-29)     float64 slab test; zero-direction
                                      (inf / NaN) cases.
  traceMap (map.glsl:83-168)         (1) hand-derived known answers from the
                                      reference code: step counts, faces,
                                      hit_pos, the 0.001 zero-direction
                                      substitution, the corner skip at an
                                      exact tie; (2) brute-force first solid
                                      voxel along the exact line on random
                                      tiny grids; (3) an independent
                                      Amanatides-Woo walk in float64 whose
                                      tDelta is 0.999/|d| on negative axes
                                      (the reference's re-anchoring drift),
                                      cell sequence equal on every ray whose
                                      crossings are not within 1e-4 of a tie.
  traceEntities (map.glsl:172-201)   closed-form slab tests incl. the
                                      reference's missing t >= 0 check.
  SkyDome2 (camera.glsl:11-19)       float64 closed form.
  composite (blit.fragment.glsl:30)  recomputed from the oracle's G-buffer.
  present (blit.fragment.glsl:32-35) crosshair pixel set = float64 brute
                                      force of |texPos - 0.5| <= 0.002 at
                                      pixel centres; vignette = float64
                                      closed form (and centre symmetry);
                                      rgba8 store: clamp, round-to-nearest
                                      known answers, R in the low byte.
  Lambert shading + FP64 gradient    float64 recomputation; central finite
                                      differences of the FP64 loss.
  chunk table / block pool / atlas   a brickmap world built like
  (map.glsl:31-60, voxel.zig:25-131) VoxelBrickmap.set renders exactly like
                                      the dense world it encodes.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_ffi as O
from vt import camera as vcam
from vt import scenes as S

F32 = np.float32


def f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------------------
# raygen
# ---------------------------------------------------------------------------

def raygen64(cam, W, H, px, py):
    th = math.tan(f32(cam.fov) / 2.0)
    u = (px / W * 2 - 1) * th
    v = (py / H * 2 - 1) * (H / W) * th
    M = np.array(cam.view, dtype=np.float64).reshape(4, 4)  # row j = column j
    r = M[0] * u + M[1] * v + M[2] + M[3]
    r = r / np.linalg.norm(r)
    return r[:3]


def test_raygen_matches_closed_form():
    rng = np.random.default_rng(1)
    L = O.lib()
    for _ in range(40):
        eye = rng.uniform(-50, 150, 3)
        tgt = rng.uniform(0, 128, 3)
        cam = vcam.look_at(eye, tgt, fov_deg=rng.uniform(30, 120))
        W, H = int(rng.integers(16, 400)), int(rng.integers(16, 400))
        oc = O.camera(cam)
        th = L.or_tan_half_fov(cam.fov)
        for _ in range(10):
            px, py = int(rng.integers(0, W)), int(rng.integers(0, H))
            o = (C.c_float * 3)()
            d = (C.c_float * 3)()
            L.or_raygen(C.byref(oc), th, W, H, px, py, o, d)
            ref = raygen64(cam, W, H, px, py)
            assert np.allclose(np.array(d[:]), ref, atol=2e-6), (W, H, px, py)
            assert list(o) == [f32(x) for x in cam.position[:3]]


def test_raygen_centre_pixel_four_component_normalize():
    # identity C_view: centre pixel direction is normalize((0,0,1,1)).xyz
    cam = O.OrCamera()
    cam.view[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]
    cam.fov = f32(math.pi / 2)
    L = O.lib()
    o = (C.c_float * 3)()
    d = (C.c_float * 3)()
    L.or_raygen(C.byref(cam), L.or_tan_half_fov(cam.fov), 64, 32, 32, 16, o, d)
    assert list(d) == [0.0, 0.0, f32(1 / math.sqrt(2))]
    # pixel 0: uv = (-1, -H/W) * tan(fov/2) with no +0.5 pixel centre
    L.or_raygen(C.byref(cam), L.or_tan_half_fov(cam.fov), 64, 32, 0, 0, o, d)
    r = np.array([-1.0, -0.5, 1.0, 1.0])
    assert np.allclose(d[:], (r / np.linalg.norm(r))[:3], atol=1e-7)


# ---------------------------------------------------------------------------
# intersectAABB and the entry point
# ---------------------------------------------------------------------------

def aabb(o, d, lo, hi):
    L = O.lib()
    tn, tf = C.c_float(), C.c_float()
    L.or_intersect_aabb(O.fvec(o), O.fvec(d), O.fvec(lo), O.fvec(hi), C.byref(tn), C.byref(tf))
    return tn.value, tf.value


def test_aabb_matches_slab_test():
    rng = np.random.default_rng(2)
    for _ in range(500):
        o = rng.uniform(-20, 40, 3).astype(F32)
        d = rng.normal(size=3).astype(F32)
        lo, hi = np.zeros(3), np.full(3, 16.0)
        tn, tf = aabb(o, d, lo, hi)
        o64, d64 = o.astype(np.float64), d.astype(np.float64)
        a, b = (lo - o64) / d64, (hi - o64) / d64
        assert math.isclose(tn, np.max(np.minimum(a, b)), rel_tol=1e-5, abs_tol=1e-5)
        assert math.isclose(tf, np.min(np.maximum(a, b)), rel_tol=1e-5, abs_tol=1e-5)


def test_aabb_zero_direction_components():
    # parallel to x inside the x slab: x contributes (-inf, +inf)
    tn, tf = aabb((3, -5, 4), (0, 1, 0), (0, 0, 0), (8, 8, 8))
    assert (tn, tf) == (5.0, 13.0)
    # origin on the x = 0 plane with d.x = 0: tMin.x = 0/0 = NaN and
    # tMax.x = +inf; min/max keep the non-NaN operand, so both t1.x and t2.x
    # are +inf and the ray reads as a miss (tNear = +inf)
    tn, tf = aabb((0, -5, 4), (0, 1, 0), (0, 0, 0), (8, 8, 8))
    assert (tn, tf) == (math.inf, 13.0)
    # parallel and outside the x slab: (0-9)/0 = (8-9)/0 = -inf, so
    # tFar = -inf < tNear (a miss)
    tn, tf = aabb((9, -5, 4), (0, 1, 0), (0, 0, 0), (8, 8, 8))
    assert tn == 5.0 and tf == -math.inf


def test_entry_origin():
    L = O.lib()
    out = (C.c_float * 3)()
    L.or_entry_origin(O.fvec((4.5, 20.0, 4.5)), O.fvec((0.0, -1.0, 0.0)), 16, out)
    # tNear = 4 -> (4.5, 16, 4.5) - 0.001 on every component
    assert list(out) == [f32(4.5 - 0.001), f32(f32(16.0) - f32(0.001)), f32(4.5 - 0.001)]
    L.or_entry_origin(O.fvec((4.5, 8.0, 4.5)), O.fvec((0.0, -1.0, 0.0)), 16, out)
    assert list(out) == [f32(4.5 - 0.001), f32(8.0 - f32(0.001)), f32(4.5 - 0.001)]


# ---------------------------------------------------------------------------
# traceMap: hand-derived known answers
# ---------------------------------------------------------------------------

def world(dim, solid=(), mat=1):
    g = np.zeros((dim, dim, dim), np.uint8)
    for (x, y, z) in solid:
        g[z, y, x] = mat
    return O.DenseWorld(g, S.palette())


def test_empty_world_steps_and_exit():
    # (0.5,0.5,0.5)+t(1,0,0): one lookup per block, 4 blocks, then the bounds test
    h, path = O.trace(world(4), (0.5, 0.5, 0.5), (1, 0, 0), 100, path_cap=16)
    assert h.data == 0 and h.steps == 4 and h.face == 0
    assert list(h.hit_pos) == [-1.0, -1.0, -1.0]
    # iteration 0 looks up pos = (4,4,4) (stepSize 0), then block-level pos
    assert [tuple(p >> 3) for p in path] == [(0, 0, 0), (1, 0, 0), (2, 0, 0), (3, 0, 0)]
    # the cap counts iterations
    assert O.trace(world(4), (0.5, 0.5, 0.5), (1, 0, 0), 2)[0].steps == 2
    assert O.trace(world(4), (0.5, 0.5, 0.5), (1, 0, 0), 0)[0].steps == 0
    # negative direction: anchored at 0.999, still one lookup per block
    h, path = O.trace(world(4), (3.5, 0.5, 0.5), (-1, 0, 0), 100, path_cap=16)
    assert h.steps == 4 and [tuple(p >> 3) for p in path] == [(3, 0, 0), (2, 0, 0), (1, 0, 0),
                                                            (0, 0, 0)]


def test_first_hit_face_and_position():
    h, _ = O.trace(world(4, [(2, 0, 0)]), (0.5, 0.5, 0.5), (1, 0, 0), 100)
    assert h.data == S.palette()[1] and h.steps == 3 and h.face == 1
    assert list(h.block) == [2, 0, 0] and list(h.normal) == [-1.0, 0.0, 0.0]
    # hit_pos = gridsCoords + within in x8 units: x on the face, y/z drifted
    # by the 0.001 substitution for the zero components over t = 12 (x8)
    assert h.hit_pos[0] == 16.0
    assert h.hit_pos[1] == f32(4.0 + f32(0.001) * 4 + f32(0.001) * 8)
    # from the other side (-x): face 2, normal +x
    h, _ = O.trace(world(4, [(1, 0, 0)]), (3.5, 0.5, 0.5), (-1, 0, 0), 100)
    assert h.face == 2 and list(h.block) == [1, 0, 0] and h.steps == 3
    # -y / +z faces
    h, _ = O.trace(world(4, [(0, 0, 0)]), (0.5, 3.5, 0.5), (0, -1, 0), 100)
    assert h.face == 4 and list(h.normal) == [0.0, 1.0, 0.0]
    h, _ = O.trace(world(4, [(0, 0, 3)]), (0.5, 0.5, 0.5), (0, 0, 1), 100)
    assert h.face == 5 and list(h.normal) == [0.0, 0.0, -1.0]


def test_start_inside_solid_reports_x_face():
    # minIdx starts at 0, so a hit on iteration 0 reports an x face whatever
    # the direction (map.glsl:98,120-121)
    h, _ = O.trace(world(4, [(1, 1, 1)]), (1.5, 1.5, 1.5), (0, -1, 0), 100)
    assert h.steps == 1 and h.face == 1
    h, _ = O.trace(world(4, [(1, 1, 1)]), (1.5, 1.5, 1.5), (-0.3, -1, 0), 100)
    assert h.steps == 1 and h.face == 2


def test_exact_tie_skips_the_corner_blocks():
    # d = (1, 1, 0.5) from (0.25,0.25,0.25): within = (2,2,2) after the
    # switch, t = (6, 6, 12): a tie; minIdx = y (t.x < t.y is false), so
    # within.x = 2 + 6 = 8 and the next lookup is pos >> 3 = G + (1,0,0)
    # = (1,1,0): blocks (1,0,0) and (0,1,0) are never looked up.
    h, path = O.trace(world(4, [(1, 0, 0)]), (0.25, 0.25, 0.25), (1, 1, 0.5), 100, 16)
    blocks = [tuple(p >> 3) for p in path]
    assert blocks[:3] == [(0, 0, 0), (1, 1, 0), (1, 1, 0)]
    assert (1, 0, 0) not in blocks and (0, 1, 0) not in blocks
    assert h.data == 0
    # a solid (1,1,0) is hit on the tie iteration with the y face
    h, _ = O.trace(world(4, [(1, 1, 0)]), (0.25, 0.25, 0.25), (1, 1, 0.5), 100)
    assert h.steps == 2 and h.face == 3 and list(h.block) == [1, 1, 0]
    assert list(h.hit_pos) == [8.0, 8.0, 5.0]


ANCHOR = f32(f32(8.0) * f32(0.999))  # within after a negative crossing (map.glsl:162)


def test_exact_two_axis_tie_on_negative_axes():
    # d = (-1, -1, -0.5) from (3.75,3.75,3.75) in a 4^3 world: o*8 = 30, so
    # iteration 0 looks up block (3,3,3) and switches with within = (6,6,6),
    # G = (24,24,24).  t = (0 - 6) * (1/d) = (6, 6, 12): a tie of x and y;
    # minIdx = y (t.x < t.y is false), within += d*6 = (0, 0, 3), then
    # within.y = 8*0.999.  Iteration 1 looks up pos = (24, 16+7, 24+3) ->
    # block (3,2,3): unlike a positive tie, one corner block IS looked up,
    # and the x crossing follows as its own step with t.x = (0-0)*(-1) = -0
    # (the smallest t), moving nothing but x.  Iteration 2: block (2,2,3).
    # Then t = (7.992, 7.992, 6): z; within = (1.992, 1.992, 0) -> anchor
    # on z; block (2,2,2); t = (1.992, 1.992, 15.98): y (x/y tie again),
    # block (2,1,2); t.x = -0: x, block (1,1,2).
    h, path = O.trace(world(4), (3.75, 3.75, 3.75), (-1, -1, -0.5), 100, 16)
    blocks = [tuple(p >> 3) for p in path]
    assert blocks[:6] == [(3, 3, 3), (3, 2, 3), (2, 2, 3), (2, 2, 2), (2, 1, 2), (1, 1, 2)]
    assert (2, 3, 3) not in blocks
    # the corner block (3,2,3) is hit with the y face (normal +y), hit_pos =
    # G + within = (24 + 0, 16 + 8*0.999, 24 + 3)
    h, _ = O.trace(world(4, [(3, 2, 3)]), (3.75, 3.75, 3.75), (-1, -1, -0.5), 100)
    assert h.steps == 2 and h.face == 4 and list(h.block) == [3, 2, 3]
    assert list(h.hit_pos) == [24.0, f32(f32(16.0) + ANCHOR), 27.0]
    assert h.tie_margin == 0.0
    # block (2,2,3) after the t = -0 step: x face (normal +x), within.x
    # re-anchored, y still the anchor
    h, _ = O.trace(world(4, [(2, 2, 3)]), (3.75, 3.75, 3.75), (-1, -1, -0.5), 100)
    assert h.steps == 3 and h.face == 2 and list(h.block) == [2, 2, 3]
    assert list(h.hit_pos) == [f32(f32(16.0) + ANCHOR), f32(f32(16.0) + ANCHOR), 27.0]


def test_exact_three_axis_tie_on_negative_axes():
    # d = (-1,-1,-1) from (3.75,3.75,3.75): t = (6,6,6), minIdx = z (both
    # strict comparisons fail); within = (0,0,0) then z anchored.  The next
    # two steps have t = -0 on x and y: y first (t.x < t.y is false for
    # -0 < -0, t.y < t.z holds), then x.  Every diagonal move therefore
    # looks up the z-neighbour and the yz-neighbour before the diagonal
    # block: z, y, x order, three lookups per diagonal step.
    h, path = O.trace(world(4), (3.75, 3.75, 3.75), (-1, -1, -1), 100, 16)
    blocks = [tuple(p >> 3) for p in path]
    assert blocks == [(3, 3, 3), (3, 3, 2), (3, 2, 2), (2, 2, 2), (2, 2, 1), (2, 1, 1),
                      (1, 1, 1), (1, 1, 0), (1, 0, 0), (0, 0, 0)]
    assert h.data == 0 and h.steps == 10
    h, _ = O.trace(world(4, [(3, 2, 2)]), (3.75, 3.75, 3.75), (-1, -1, -1), 100)
    assert h.steps == 3 and h.face == 4 and list(h.block) == [3, 2, 2]
    a = f32(f32(16.0) + ANCHOR)
    assert list(h.hit_pos) == [24.0, a, a]
    h, _ = O.trace(world(4, [(2, 2, 2)]), (3.75, 3.75, 3.75), (-1, -1, -1), 100)
    assert h.steps == 4 and h.face == 2 and list(h.hit_pos) == [a, a, a]


def test_outside_start_is_a_miss_without_lookups():
    h, _ = O.trace(world(4), (-1.0, 0.5, 0.5), (1, 0, 0), 100)
    assert h.steps == 0 and h.data == 0


# ---------------------------------------------------------------------------
# traceMap vs brute force and vs an independent Amanatides-Woo walk
# ---------------------------------------------------------------------------

def aw_walk(o, d, dim, max_cells):
    """Cell sequence of the line o + t d (voxel units), float64, tDelta on
    negative axes 0.999/|d| (the reference re-anchors a negative crossing at
    0.999 of the new cell).  Returns (cells, min gap between the two smallest
    tMax at every step)."""
    d = np.array([x if x != 0 else f32(0.001) for x in d], dtype=np.float64)
    o = np.asarray(o, dtype=np.float64)
    cell = np.floor(o).astype(int)
    step = np.where(d > 0, 1, -1)
    frac = o - cell
    t_max = np.where(d > 0, (1 - frac) / d, frac / -d)
    t_delta = np.where(d > 0, 1 / d, 0.999 / -d)
    cells, gap = [], math.inf
    while len(cells) < max_cells and np.all((cell >= 0) & (cell < dim)):
        cells.append(tuple(cell))
        s = np.sort(t_max)
        gap = min(gap, s[1] - s[0])
        m = int(np.argmin(t_max))
        cell[m] += step[m]
        t_max[m] += t_delta[m]
    return cells, gap


def first_hit_bruteforce(grid, o, d):
    """Exact first solid voxel along o + t d, t >= 0 (float64 slab tests).

    Returns (voxel or None, sep) where sep is the smaller of the gap between
    the two earliest entries and the closest graze (|tFar - tNear| of any
    solid box ahead): below ~0.01 the reference's 0.001-per-crossing drift
    can legitimately decide differently."""
    d = np.array([x if x != 0 else 1e-30 for x in d], np.float64)
    zs, ys, xs = np.nonzero(grid)
    if len(xs) == 0:
        return None, math.inf
    lo = np.stack([xs, ys, zs], 1).astype(np.float64)
    a, b = (lo - o) / d, (lo + 1 - o) / d
    tn, tf = np.max(np.minimum(a, b), 1), np.min(np.maximum(a, b), 1)
    ahead = tf >= -0.05
    graze = np.min(np.abs(tf - tn)[ahead]) if ahead.any() else math.inf
    ok = (tf >= np.maximum(tn, 0))
    if not ok.any():
        return None, graze
    tnc = np.where(ok, np.maximum(tn, 0), np.inf)
    order = np.argsort(tnc)
    sep = tnc[order[1]] - tnc[order[0]] if ok.sum() > 1 else math.inf
    i = order[0]
    return (int(xs[i]), int(ys[i]), int(zs[i])), min(sep, graze)


def test_trace_matches_amanatides_woo_with_drift():
    rng = np.random.default_rng(3)
    dim = 16
    w = world(dim)
    checked = 0
    for _ in range(3000):
        o = rng.uniform(0.05, dim - 0.05, 3).astype(F32)
        d = rng.normal(size=3).astype(F32)
        d /= np.float32(np.linalg.norm(d))
        h, path = O.trace(w, o, d, 400, path_cap=400)
        cells, gap = aw_walk(o, d, dim, 400)
        if gap < 1e-4 or h.tie_margin < 1e-4:
            continue
        got = [tuple(p >> 3) for p in path]
        assert got == cells, (o, d)
        assert h.steps == len(cells) and h.data == 0
        checked += 1
    assert checked > 2500


def test_trace_matches_bruteforce_first_hit():
    rng = np.random.default_rng(4)
    dim = 8
    checked = 0
    for trial in range(300):
        grid = (rng.random((dim, dim, dim)) < 0.08).astype(np.uint8) * 3
        w = O.DenseWorld(grid, S.palette())
        for _ in range(20):
            o = rng.uniform(0.05, dim - 0.05, 3).astype(F32)
            d = rng.normal(size=3).astype(F32)
            h, _ = O.trace(w, o, d, 200)
            exp, sep = first_hit_bruteforce(grid, o.astype(np.float64), d.astype(np.float64))
            if sep < 0.05 or h.tie_margin < 1e-3 or h.steps > 12:
                continue  # near-ties and long paths (drift) are left to the AW test
            if exp is None:
                assert h.data == 0
            else:
                assert h.data != 0 and tuple(h.block) == exp, (trial, o, d, exp)
                assert grid[exp[2], exp[1], exp[0]] != 0
            checked += 1
    assert checked > 2000


def test_hit_face_is_the_entered_face():
    rng = np.random.default_rng(5)
    dim = 8
    n = 0
    for _ in range(200):
        grid = (rng.random((dim, dim, dim)) < 0.1).astype(np.uint8)
        w = O.DenseWorld(grid, S.palette())
        o = rng.uniform(0.05, dim - 0.05, 3).astype(F32)
        if grid[int(o[2]), int(o[1]), int(o[0])]:
            continue
        d = rng.normal(size=3).astype(F32)
        h, path = O.trace(w, o, d, 200, path_cap=200)
        if h.data == 0 or h.tie_margin < 1e-3 or len(path) < 2:
            continue
        prev, cur = np.array(path[-2]) >> 3, np.array(path[-1]) >> 3
        delta = cur - prev
        if np.abs(delta).sum() != 1:
            continue
        axis = int(np.argmax(np.abs(delta)))
        expected = 2 * axis + (1 if delta[axis] > 0 else 2)
        assert h.face == expected
        n += 1
    assert n > 50


# ---------------------------------------------------------------------------
# traceEntities
# ---------------------------------------------------------------------------

def ent(o, d, maxd):
    return O.lib().or_trace_entities(O.fvec(o), O.fvec(d), f32(maxd))


def test_entities_closed_form():
    # straight down onto the box at (256,21,256)
    assert ent((256.5, 30.0, 256.5), (0, -1, 0), 1000.0) == 1
    # culled: every box farther than maxDistance
    assert ent((256.5, 30.0, 256.5), (0, -1, 0), 5.0) == 0
    # the reference tests tFar >= tNear without t >= 0: a box BEHIND the
    # origin occludes (map.glsl:192)
    assert ent((256.5, 30.0, 256.5), (0, 1, 0), 1000.0) == 1
    # a line that misses every box
    assert ent((200.0, 30.0, 200.0), (0, 1, 0), 1000.0) == 0
    # (251,21,259) box from the side along +x
    assert ent((240.0, 21.5, 259.5), (1, 0, 0), 1000.0) == 1


# ---------------------------------------------------------------------------
# SkyDome2, composite, Lambert
# ---------------------------------------------------------------------------

def sky64(sun, rd):
    sun, rd = np.asarray(sun, float), np.asarray(rd, float)
    s = min(max(np.dot(sun / np.linalg.norm(sun), rd / np.linalg.norm(rd)), 0.0), 1.2)
    col = np.array([0.6, 0.71, 0.75]) - rd[1] * 0.2 * np.array([1.0, 0.5, 1.0]) + 0.075
    col += 0.4 * np.array([1.0, 0.6, 0.1]) * s ** 8
    col += np.array([0.2, 0.08, 0.04]) * s ** 3
    return col


def test_sky_closed_form():
    rng = np.random.default_rng(6)
    L = O.lib()
    out = (C.c_float * 3)()
    for _ in range(200):
        sun, rd = rng.normal(size=3), rng.normal(size=3)
        L.or_sky(O.fvec(sun), O.fvec(rd), out)
        assert np.allclose(out[:], sky64(np.float32(sun), np.float32(rd)), atol=2e-6)


@pytest.fixture(scope="module")
def c0_render():
    grid = S.generate(S.spec("c0"))
    cam, W, H = S.views("c0")[0]
    w = O.DenseWorld(grid, S.palette())
    return grid, w, cam, W, H, O.render(w, cam, W, H, 512, 512, S.SUN)


def test_composite_and_gbuffer_rules(c0_render):
    grid, w, cam, W, H, r = c0_render
    hit = r["hits"][..., 0] != 0xFFFFFFFF
    col, ill, pos, nrm = r["color"], r["illumination"], r["position"], r["normal"]
    # sky pixels: position -1, normal 1, illumination 0, shaded = sky colour
    assert np.all(pos[~hit] == -1) and np.all(nrm[~hit] == 1) and np.all(ill[~hit] == 0)
    assert np.array_equal(r["shaded"][~hit], col[~hit])
    # hits: illumination = (sun, +-0.3) and shaded = color + a * sky(sun)
    sun = np.float32(S.SUN)
    assert np.all(ill[hit][:, :3] == sun)
    occ = (r["hits"][..., 2] >> 3) & 1
    assert np.all(ill[hit][:, 3] == np.where(occ[hit] == 1, np.float32(-0.3), np.float32(0.3)))
    sk = sky64(sun, sun)
    exp = col[hit][:, :3] + ill[hit][:, 3:4] * sk
    assert np.max(np.abs(r["shaded"][hit][:, :3] - exp)) < 2e-6
    # colour of a hit = palette colour of the hit voxel's material
    pal = S.palette()
    mat = (r["hits"][..., 2] >> 16) & 255
    vid = r["hits"][..., 0][hit]
    n = grid.shape[0]
    assert np.array_equal(mat[hit], grid.ravel()[vid])
    rgba = np.stack([(pal[mat[hit]] >> (8 * k)) & 255 for k in range(4)], 1) / np.float32(255)
    assert np.array_equal(col[hit], rgba.astype(np.float32))
    # position = ceil(hit_pos)/8 lies on the hit voxel's closed box
    x, y, z = vid % n, (vid // n) % n, vid // (n * n)
    p = pos[hit][:, :3]
    assert np.all((p >= np.stack([x, y, z], 1) - 1e-6) & (p <= np.stack([x, y, z], 1) + 1 + 1e-6))
    # the shadow ray count equals the hit count, every hit voxel is solid
    assert int(r["stats"][2]) == int(hit.sum())
    assert np.all(grid.ravel()[vid] != 0)


def test_unorm8_mode_quantizes(c0_render):
    grid, w, cam, W, H, _ = c0_render
    r = O.render(w, cam, W, H, 512, 512, S.SUN, flags=1)
    hit = r["hits"][..., 0] != 0xFFFFFFFF
    q = lambda x: np.floor(np.clip(x, 0, 1) * np.float32(255) + np.float32(0.5)) / np.float32(255)
    ill = r["illumination"][hit]
    assert np.all(ill[:, :3] == q(np.float32(S.SUN)).astype(np.float32))
    assert set(np.unique(ill[:, 3])) <= {np.float32(0.0), np.float32(77 / 255)}
    assert set(np.unique(r["normal"][hit][:, :3])) <= {np.float32(0.0), np.float32(1.0)}


def test_lambert_and_fp64_gradient(c0_render):
    grid, w, cam, W, H, r = c0_render
    rng = np.random.default_rng(7)
    n = grid.shape[0]
    albedo = rng.uniform(0, 1, n ** 3 * 3).astype(np.float32)
    hits = np.ascontiguousarray(r["hits"].reshape(-1, 3))
    npix = hits.shape[0]
    L = O.lib()
    out = np.zeros((npix, 4), np.float32)
    sun, amb = np.float32(S.SUN), np.float32(0.1)
    L.or_shade_lambert(hits.ctypes.data, npix, albedo.ctypes.data, O.fvec(sun), amb,
                       out.ctypes.data)
    hit = hits[:, 0] != 0xFFFFFFFF
    face = hits[:, 2] & 7
    normals = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]])
    l = sun.astype(np.float64) / np.linalg.norm(sun.astype(np.float64))
    nd = np.where(hit, normals[np.maximum(face, 1) - 1] @ l, 0)
    vis = 1 - ((hits[:, 2] >> 3) & 1)
    fct = amb + np.maximum(nd, 0) * vis
    vid = np.where(hit, hits[:, 0], 0)
    exp = albedo.reshape(-1, 3)[vid] * fct[:, None]
    assert np.allclose(out[hit, :3], exp[hit], atol=1e-6)
    assert np.all(out[~hit] == 0)

    # FP64 gradient vs central differences of the FP64 loss (a small image)
    sub = np.ascontiguousarray(hits[:4096])
    tgt = rng.uniform(0, 1, (4096, 4)).astype(np.float32)

    def loss64(alb, s):
        da = np.zeros(alb.size)
        ds = np.zeros(3)
        return L.or_lambert_grad64(sub.ctypes.data, 4096, alb.ctypes.data, O.fvec(s), amb,
                                   tgt.ctypes.data, da.ctypes.data, ds.ctypes.data), da, ds

    l0, da, ds = loss64(albedo, sun)
    used = np.unique(sub[sub[:, 0] != 0xFFFFFFFF][:, 0])
    for v in used[:8]:
        for c in range(3):
            a1, a2 = albedo.copy(), albedo.copy()
            h = np.float32(1e-2)
            a1[3 * v + c] += h
            a2[3 * v + c] -= h
            fd = (loss64(a1, sun)[0] - loss64(a2, sun)[0]) / (a1[3 * v + c] - a2[3 * v + c])
            assert math.isclose(fd, da[3 * v + c], rel_tol=1e-4, abs_tol=1e-6)
    # the loss is linear in albedo per voxel... and smooth in the sun direction
    for k in range(3):
        s1, s2 = sun.copy(), sun.copy()
        s1[k] += np.float32(1e-3)
        s2[k] -= np.float32(1e-3)
        fd = (loss64(albedo, s1)[0] - loss64(albedo, s2)[0]) / float(s1[k] - s2[k])
        assert math.isclose(fd, ds[k], rel_tol=2e-3, abs_tol=1e-3)


# ---------------------------------------------------------------------------
# chunk table / block pool / atlas layout
# ---------------------------------------------------------------------------

def brickmap_from_dense(grid, palette):
    """Reference layout built the way VoxelBrickmap.set does (voxel.zig:47-61):
    chunks in first-touch order, block value = material k, atlas slot k = a
    solid 8^3 cube of palette[k] (slot origin ((k%32)*8, (k/32%32)*8,
    (k/1024)*8), voxel.zig:100-110)."""
    dim = grid.shape[0]
    M = dim // 8
    chunks = np.zeros(M ** 3, np.uint32)
    pool = []
    zs, ys, xs = np.nonzero(grid)
    for x, y, z in zip(xs, ys, zs):
        ci = x // 8 + M * (y // 8 + (z // 8) * M)
        if chunks[ci] == 0:
            pool.append(np.zeros(512, np.uint32))
            chunks[ci] = len(pool)
        pool[chunks[ci] - 1][x % 8 + ((y % 8) + (z % 8) * 8) * 8] = grid[z, y, x]
    atlas = np.zeros((256, 256, 256), np.uint32)
    for k in range(1, 256):
        ox, oy, oz = (k % 32) * 8, (k // 32 % 32) * 8, (k // 1024) * 8
        atlas[oz:oz + 8, oy:oy + 8, ox:ox + 8] = palette[k]
    pool = np.concatenate(pool) if pool else np.zeros(0, np.uint32)
    return chunks, pool, atlas.ravel()


def test_brickmap_layout_renders_like_dense(c0_render):
    grid, w, cam, W, H, r = c0_render
    chunks, pool, atlas = brickmap_from_dense(grid, S.palette())
    bw = O.BrickmapWorld(grid.shape[0], chunks, pool, atlas)
    rb = O.render(bw, cam, W, H, 512, 512, S.SUN)
    for k in ("hits", "steps", "position", "normal", "illumination", "shaded"):
        assert np.array_equal(rb[k], r[k]), k


def test_two_level_traversal_hits_solid_subvoxels():
    """Embedding C (grid voxel = sub-voxel, 8^3 bricks as atlas models): every
    reported hit is a solid sub-voxel of the source grid, its face is on the
    side the ray came from, and nearly every pixel agrees with the dense
    traversal of the same grid (they differ only where the two levels' drift
    decides a grazing ray differently)."""
    import brickmap_util as BU
    grid = S.generate(S.spec("c0"))
    pal = S.palette()
    dim, chunks, pool, atlas = BU.embedding_c(grid, pal)
    bw = O.BrickmapWorld(dim, chunks, pool, atlas)
    cam = vcam.look_at((8.0, 13.75, -2.5), (8.0, 8.0, 8.0), 70.0)
    r = O.render(bw, cam, 192, 192, 1024, 256, S.SUN)
    h = r["hits"].reshape(-1, 3)
    hit = h[:, 0] != 0xFFFFFFFF
    assert hit.mean() > 0.3
    v, sub = h[hit, 0], (h[hit, 2] >> 4) & 511
    bx, by, bz = v % dim, (v // dim) % dim, v // (dim * dim)
    sx, sy, sz = bx * 8 + (sub & 7), by * 8 + ((sub >> 3) & 7), bz * 8 + (sub >> 6)
    assert np.all(grid[sz, sy, sx] != 0)
    # the dense traversal of the same grid with the camera scaled by 8
    camd = vcam.look_at((64.0, 110.0, -20.0), (64.0, 64.0, 64.0), 70.0)
    rd = O.render(O.DenseWorld(grid, pal), camd, 192, 192, 1024, 256, S.SUN)
    hd = rd["hits"].reshape(-1, 3)
    vd = hd[hit, 0].astype(np.int64)
    same = vd == sx + 128 * (sy + 128 * sz)
    assert same.mean() > 0.95
    # the two traversals drift by different amounts (0.001 block per negative
    # crossing vs 0.001 sub-voxel at the fine level): where they part, the
    # hit voxels are neighbours (Chebyshev distance 1) on all but grazing rays
    ok = vd != 0xFFFFFFFF
    cheb = np.maximum(np.maximum(np.abs(vd % 128 - sx), np.abs(vd // 128 % 128 - sy)),
                      np.abs(vd // 16384 - sz))
    assert np.mean(cheb[ok] <= 1) > 0.99


# ---------------------------------------------------------------------------
# present pass
# ---------------------------------------------------------------------------

def _crosshair64(W, H):
    x = (np.arange(W) + 0.5) / W - 0.5
    y = (np.arange(H) + 0.5) / H - 0.5
    d = np.sqrt(x[None, :] ** 2 + y[:, None] ** 2)
    return d, d <= 0.002


@pytest.mark.parametrize("W,H", [(1920, 1080), (256, 256), (1000, 1000), (999, 1001), (7, 5)])
def test_present_crosshair_set(W, H):
    out, _ = O.present(np.zeros((H, W, 4), np.float32))
    got = out[..., 0] > 0  # zero composite: only crosshair pixels turn grey
    d, exp = _crosshair64(W, H)
    ambiguous = np.abs(d - 0.002) < 1e-7
    assert np.array_equal(got[~ambiguous], exp[~ambiguous])
    # the crosshair colour is mix(0, (1,1,1,0.4), 0.5) times the vignette
    g = out[got][:, 3] / 0.2
    assert np.allclose(out[got][:, 0], 0.5 * g, rtol=1e-6)


def test_present_crosshair_known_sizes():
    # 1000x1000: radius 0.002 = 2 pixels, centres at offsets +-0.5, +-1.5 px
    out, _ = O.present(np.zeros((1000, 1000, 4), np.float32))
    ys, xs = np.nonzero(out[..., 0] > 0)
    assert (xs.min(), xs.max(), ys.min(), ys.max()) == (498, 501, 498, 501)
    assert len(xs) == 12  # 4x4 square minus the 4 corners (|(1.5,1.5)| > 2)
    # a 7x5 image has no pixel centre within 0.002 of the middle except (3, 2)
    out, _ = O.present(np.zeros((5, 7, 4), np.float32))
    assert list(zip(*np.nonzero(out[..., 0] > 0))) == [(2, 3)]


@pytest.mark.parametrize("W,H", [(1920, 1080), (333, 77)])
def test_present_vignette_closed_form(W, H):
    comp = np.ones((H, W, 4), np.float32)
    out, _ = O.present(comp)
    u = (np.arange(W) + 0.5) / W
    v = (np.arange(H) + 0.5) / H
    g = (u[None, :] * (1 - u[None, :]) * v[:, None] * (1 - v[:, None]) * 15.0) ** 0.18
    _, cross = _crosshair64(W, H)
    assert np.max(np.abs(out[~cross][:, 0] - g[~cross])) <= 2e-6
    # the pixel-centre mapping is symmetric about the image centre
    assert np.max(np.abs(out[..., 1] - out[::-1, ::-1, 1])) <= 2e-6
    # centre of an even image: u = v = 0.5 +- 1/(2W) -> g ~ (15/16)^0.18
    assert abs(out[H // 2, W // 2 - 3, 2] - (15 / 16) ** 0.18) < 1e-4


def test_present_rgba8_store():
    comp = np.zeros((3, 4, 4), np.float32)
    comp[0, 0] = [-1.0, 2.0, 0.5, 1.0]
    comp[0, 1] = [np.nan, 0.25, 0.75, 0.0]
    out, disp = O.present(comp)
    d = disp[0, 0]
    r, g, b, a = d & 255, (d >> 8) & 255, (d >> 16) & 255, d >> 24
    assert r == 0 and g == 255  # clamped
    exp_b = math.floor(f32(f32(out[0, 0, 2]) * 255.0) + 0.5)
    assert b == exp_b and a == math.floor(f32(out[0, 0, 3] * 255.0) + 0.5)
    # half-way cases round up: 0.5 * 255 = 127.5 -> 128
    comp2 = np.full((1, 1, 4), 0.5, np.float32)
    out2, disp2 = O.present(comp2)
    k = out2[0, 0, 0]
    assert disp2[0, 0] & 255 == int(np.floor(np.float32(k) * np.float32(255) + np.float32(0.5)))


# ---------------------------------------------------------------------------
# centre-ray pick (terrain_edit.comp.glsl:11-24)
# ---------------------------------------------------------------------------

def test_pick_is_the_centre_pixel_primary_ray(c0_render):
    """rayUV = 0 is exactly the primary ray of pixel (W/2, H/2) of an even
    image (uv = (W/2)/W*2-1 = 0), so the pick must report that pixel's hit
    when traced with the pick's 64-step cap."""
    grid, w, _, _, _, _ = c0_render
    rng = np.random.default_rng(11)
    W = H = 64
    hits = 0
    for k in range(24):
        eye = (rng.uniform(0, 128), rng.uniform(70, 130), rng.uniform(-40, 0))
        tgt = (rng.uniform(20, 108), rng.uniform(20, 60), rng.uniform(20, 108))
        cam = vcam.look_at(eye, tgt, 70.0)
        p = O.pick(w, cam, 64)
        r = O.render(w, cam, W, H, 64, 64, S.SUN, rows=(H // 2, H // 2 + 1), gbuffer=True)
        rec = r["hits"][H // 2, W // 2]
        if rec[0] == 0xFFFFFFFF:
            assert p.data == 0 and p.face == 0 and list(p.hit_pos) == [-1, -1, -1]
            assert list(p.place) == [-1, -1, -1] and list(p.remove) == [-1, -1, -1]
            continue
        hits += 1
        vid = int(rec[0])
        block = (vid % 128, vid // 128 % 128, vid // 16384)
        assert p.face == (rec[2] & 7) and p.data == S.palette()[grid[block[2], block[1], block[0]]]
        assert tuple(p.remove) == block
        n = np.array(p.normal, np.int32)
        assert tuple(p.place) == tuple(np.array(block) + n)  # the empty neighbour cell
        assert grid[p.place[2], p.place[1], p.place[0]] == 0 if min(p.place) >= 0 and max(p.place) < 128 else True
        # the G-buffer position is the ceil of the same hit position / 8
        pos = r["position"][H // 2, W // 2, :3]
        assert np.array_equal(np.ceil(np.array(p.hit_pos, np.float32)) / np.float32(8), pos)
    assert hits >= 12


def test_pick_known_answer():
    """A single voxel seen from above: face +y, remove = the voxel, place =
    the cell on top of it; a camera looking away misses."""
    grid = np.zeros((16, 16, 16), np.uint8)
    grid[7, 3, 5] = 2  # (x, y, z) = (5, 3, 7)
    pal = S.palette()
    w = O.DenseWorld(grid, pal)
    cam = vcam.look_at((5.5, 12.0, 7.4), (5.5, 0.0, 7.6), 70.0)
    p = O.pick(w, cam, 64)
    assert p.data == pal[2] and p.face == 4 and list(p.normal) == [0, 1, 0]
    assert tuple(p.remove) == (5, 3, 7) and tuple(p.place) == (5, 4, 7)
    # top face of block y=3 (sub-voxel y 32) as the reference leaves it after a
    # negative crossing at the block level: grid 24 + within (1 << 3) * 0.999
    assert p.hit_pos[1] == f32(24.0 + f32(8.0 * f32(0.999)))
    away = vcam.look_at((5.5, 12.0, 7.4), (5.5, 30.0, 7.6), 70.0)
    q = O.pick(w, away, 64)
    assert q.data == 0 and tuple(q.place) == (-1, -1, -1)
