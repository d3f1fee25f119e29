"""Pins the CPU restatement of the reference world generator (procgen.zig).

FastNoiseLite (through znoise, a build.zig.zon dependency that is not
vendored) is absent here and the reference ships no terrain, so the noise is
pinned by what the mathematics fixes, not by retyping it:

  gradient table      unit vectors; the 6 distinct magnitudes equal the
                      library's published 15-digit table entries; pairs
                      0..119 repeat with period 24 and step 15 degrees.
  simplex octave      zero at every lattice point of the skewed grid; equal
                      (1e-5) to an independent float64 evaluation as a sum
                      over ALL lattice vertices of the radial kernel
                      (0.5 - r^2)^4 (g . d) -- no containing-triangle
                      selection, no far-corner shortcut -- which pins the
                      skew/unskew constants, the corner offsets and which
                      hash goes with which corner.
  fBm                 the same float64 evaluation composed with the
                      frequency, the skew and 3 octaves (bounding 1/1.75,
                      gain 0.5, lacunarity 2, seeds 1337..1339); |fbm| <= 1.

The generator itself is checked against the reference's code as read:
the base layer's chunk allocation order, the material range of every
layer of every column, the decorations and tree shapes, and a literal
replay of the LCG over the first columns in Python integers.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O

PX, PY = 501125321, 1136930381
SOLID = 1 << 28


def grads():
    g = np.zeros(256, np.float32)
    O.lib().or_fnl_gradients(g.ctypes.data)
    return g


def test_gradient_table():
    g = grads().astype(np.float64).reshape(128, 2)
    assert np.allclose(np.hypot(g[:, 0], g[:, 1]), 1.0, atol=1e-7)
    published = np.float32([0.130526192220052, 0.38268343236509, 0.608761429008721,
                            0.793353340291235, 0.923879532511287, 0.99144486137381])
    assert np.array_equal(np.unique(np.abs(grads())), np.sort(published))
    ang = np.degrees(np.arctan2(g[:, 0], g[:, 1])) % 360.0
    assert np.allclose(ang[:120], (7.5 + 15.0 * (np.arange(120) % 24)) % 360.0, atol=1e-5)
    assert np.allclose(ang[120:], 22.5 + 45.0 * np.arange(8), atol=1e-5)


def _hash_pair(seed, vi, vj):
    """Gradient pair of lattice vertex (vi, vj): FastNoiseLite's _fnlHash2D
    then hash ^= hash >> 15 (arithmetic), pair (hash & 254) / 2."""
    h = ((seed & 0xFFFFFFFF) ^ ((vi * PX) & 0xFFFFFFFF) ^ ((vj * PY) & 0xFFFFFFFF))
    h = (h * 0x27D4EB2D) & 0xFFFFFFFF
    hs = h - (1 << 32) if h >= (1 << 31) else h
    h ^= (hs >> 15) & 0xFFFFFFFF
    return (h & 254) >> 1


def simplex64(g, seed, X, Y):
    """Sum over all lattice vertices near the (skewed) point, float64."""
    G2 = (3.0 - math.sqrt(3.0)) / 6.0
    i0, j0 = math.floor(X), math.floor(Y)
    total = 0.0
    for vi in range(i0 - 1, i0 + 3):
        for vj in range(j0 - 1, j0 + 3):
            sx, sy = X - vi, Y - vj
            t = (sx + sy) * G2
            dx, dy = sx - t, sy - t
            r = 0.5 - dx * dx - dy * dy
            if r > 0:
                p = _hash_pair(seed, vi, vj)
                total += r ** 4 * (dx * g[p, 0] + dy * g[p, 1])
    return total * 99.83685446303647


def fbm64(g, x, y):
    F2 = 0.5 * (math.sqrt(3.0) - 1.0)
    x, y = x * 0.01, y * 0.01
    t = (x + y) * F2
    x, y = x + t, y + t
    s, amp = 0.0, 1.0 / 1.75
    for o in range(3):
        s += simplex64(g, 1337 + o, x, y) * amp
        x, y, amp = x * 2.0, y * 2.0, amp * 0.5
    return s


def test_simplex_zero_on_lattice():
    f = O.lib().or_fnl_simplex2
    for seed in (0, 1337, -7):
        for i in range(-12, 13, 3):
            for j in range(-12, 13, 2):
                assert f(seed, float(i), float(j)) == 0.0


def test_simplex_equals_vertex_sum():
    g = grads().astype(np.float64).reshape(128, 2)
    rng = np.random.default_rng(3)
    f = O.lib().or_fnl_simplex2
    worst = 0.0
    for X, Y in rng.uniform(-40.0, 40.0, (3000, 2)).astype(np.float32):
        for seed in (1337, 1338):
            a = f(seed, float(X), float(Y))
            b = simplex64(g, seed, float(X), float(Y))
            worst = max(worst, abs(a - b))
    assert worst <= 1e-5, worst


def test_fbm_equals_vertex_sum_and_is_bounded():
    g = grads().astype(np.float64).reshape(128, 2)
    f = O.lib().or_fnl_fbm2
    rng = np.random.default_rng(4)
    pts = [(x / 10.0, z / 10.0) for x in range(0, 1024, 37) for z in range(0, 1024, 41)]
    pts += [tuple(p) for p in rng.uniform(-300.0, 300.0, (500, 2))]
    worst, top = 0.0, 0.0
    for x, y in pts:
        x, y = float(np.float32(x)), float(np.float32(y))
        a = f(x, y)
        worst = max(worst, abs(a - fbm64(g, x, y)))
        top = max(top, abs(a))
    assert worst <= 1e-5, worst
    assert 0.3 < top <= 1.0


def decode(dim, chunks, pool, ymax):
    """vol[x, y, z] for y < ymax from the chunk table and pool."""
    M = dim // 8
    x, y, z = np.meshgrid(np.arange(dim), np.arange(ymax), np.arange(dim), indexing="ij")
    e = chunks[(x >> 3) + M * ((y >> 3) + M * (z >> 3))].astype(np.int64)
    a = (e - 1) * 512 + (x & 7) + 8 * ((y & 7) + 8 * (z & 7))
    return np.where(e > 0, pool[np.maximum(a, 0)], 0).astype(np.uint32)


def gen512_data():
    rc, chunks, pool = O.procgen(512)
    assert rc == 0
    H = O.procgen_heights(512)
    return chunks, pool, H, decode(512, chunks, pool, 64)


@pytest.fixture(scope="module")
def gen512():
    return gen512_data()


def test_base_layer_allocation_order(gen512):
    chunks, pool, H, vol = gen512
    M = 64
    cx, cz = np.meshgrid(np.arange(M), np.arange(M), indexing="ij")
    first = 2 * (cx * M + cz) + 1
    assert np.array_equal(chunks[cx + M * (0 + M * cz)], first)
    assert np.array_equal(chunks[cx + M * (1 + M * cz)], first + 1)
    assert pool.size % 512 == 0 and chunks.max() == pool.size // 512


def test_column_layers(gen512):
    chunks, pool, H, vol = gen512
    assert H.max() < 60 and (H > 16).mean() > 0.3 and (H <= 16).mean() > 0.1
    y = np.arange(64)[None, :, None]
    h = H[:, None, :].astype(np.int64)
    ty, solid = vol & 0xFFFFFFF, vol >> 28
    low = y < np.minimum(h, 16)
    assert np.all(solid[low] == 1) and np.all((ty[low] >= 25) & (ty[low] <= 27))
    mid = (y >= 16) & (y < h - 1)
    assert np.all(solid[mid] == 1) and np.all((ty[mid] >= 21) & (ty[mid] <= 23))
    top = (y == h - 1) & (h > 16)
    assert np.all(solid[top] == 1) and np.all(ty[top] <= 5)
    base = (y >= h) & (y < 16)
    assert np.all(vol[base] == (13 | SOLID))
    # each terrain material of a layer is drawn uniformly from its 3 or 6 values
    for m, lo, n in ((low, 25, 3), (mid, 21, 3), (top, 0, 6)):
        c = np.bincount(ty[m] - lo, minlength=n)
        assert c.min() > 0.9 * c.mean()


def test_decorations_and_trees(gen512):
    chunks, pool, H, vol = gen512
    h = H.astype(np.int64)
    xs, zs = np.nonzero(h > 16)
    deco = vol[xs, h[xs, zs], zs]
    grass = (deco >= 7) & (deco <= 11)  # no solid bit
    dirt = deco == (12 | SOLID)
    leaves_or_trunk = ((deco & 0xFFFFFFF) >= 14) & (deco >> 28 == 1)
    assert np.all(grass | dirt | leaves_or_trunk | (deco == 0))
    assert abs(grass.mean() - 0.2 * (1 - 1 / 71)) < 0.01
    assert abs(dirt.mean() - 1 / 71) < 0.004
    # trees: a trunk column (x+1, z+1) of 4..7 blocks of 14..16 on top of
    # the planting column's height, 3x3x3 leaves (18..19) above it
    above = np.arange(64)[None, :, None] >= np.maximum(h[:, None, :] + 1, 16)
    treev = vol[above]
    assert np.all((treev == 0) | np.isin(treev & 0xFFFFFFF, [14, 15, 16, 18, 19]))
    trunks = 0
    ok = (xs > 5) & (xs < 500) & (zs > 5) & (zs < 500)
    xs, zs = xs[ok], zs[ok]
    # a trunk of at least 4 blocks covers (x+1, vh+3, z+1)
    cand = np.isin(vol[xs + 1, np.minimum(h[xs, zs] + 3, 63), zs + 1] & 0xFFFFFFF, [14, 16])
    for x, z in zip(xs[cand], zs[cand]):
        y = h[x, z]
        for th in range(4, 8):
            # the trunk's lower blocks may lie under the terrain of column
            # (x+1, z+1), which the generator visits later
            trunk = vol[x + 1, max(y, H[x + 1, z + 1]):y + th, z + 1] & 0xFFFFFFF
            leaves = vol[x:x + 3, y + th:y + th + 3, z:z + 3] & 0xFFFFFFF
            if np.all(np.isin(trunk, [14, 15, 16])) and np.all(np.isin(leaves, [18, 19])):
                trunks += 1
                break
    assert trunks > 100  # about 1 in 420 of ~176k eligible columns


def _lcg(s):
    s = (s * 1103515245 + 12345) & 0xFFFFFFFF
    return s, s


def test_lcg_replay_first_rows(gen512):
    """procgen.zig:10-52 replayed literally over the rows x = 0..5 (which
    plant no trees: x > 5 is required, and trees only reach x..x+2)."""
    chunks, pool, H, vol = gen512
    s = 0x46AE4F
    tall = 0
    for x in range(6):
        for z in range(512):
            vh = int(H[x, z])
            exp = np.zeros(64, np.int64)
            exp[:16] = 13 | SOLID
            for hh in range(vh):
                s, r = _lcg(s)
                exp[hh] = (21 + r % 3) | SOLID
                if hh <= 15:
                    s, r = _lcg(s)
                    exp[hh] = (25 + r % 3) | SOLID
                elif hh == vh - 1 and hh > 15:
                    s, r = _lcg(s)
                    exp[hh] = (r % 6) | SOLID
            if vh > 16:
                tall += 1
                s, r = _lcg(s)
                if r % 5 == 0:
                    s, r = _lcg(s)
                    exp[vh] = 7 + r % 5
                s, r = _lcg(s)
                if r % 71 == 0:
                    exp[vh] = 12 | SOLID
                s, r = _lcg(s)  # the tree draw (no tree in these rows)
            s, _ = _lcg(s)
            assert np.array_equal(vol[x, :, z], exp), (x, z)
    assert tall > 500
