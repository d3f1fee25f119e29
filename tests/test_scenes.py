"""Benchmark world generators: pinned bytes and the structure the workload
descriptions state (terrain layering, heights, blocks, sprite fill)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
from vt import scenes as S

MANIFEST = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "manifest.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["c0", "c3", "c1"])
def test_world_bytes_pinned(name):
    g = S.generate(S.spec(name))
    m = MANIFEST["worlds"][name]
    assert g.shape == (m["dim"],) * 3
    assert sha(g) == m["sha256"]


def test_generator_bytes_do_not_depend_on_threads():
    s = S.spec("c0", n_sprites=200)
    assert np.array_equal(S.generate(s, threads=1), S.generate(s, threads=5))


def test_generator_is_deterministic_and_seeded():
    a = S.generate(S.spec("c0"))
    b = S.generate(S.spec("c0"))
    c = S.generate(S.spec("c0", seed=S.MASTER_SEED + 1))
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_terrain_layering_and_heights():
    s = S.spec("c0", n_blocks=0)
    g = S.generate(s)
    n = g.shape[0]
    for (x, z) in [(0, 0), (17, 99), (64, 64), (127, 3), (80, 127)]:
        h = S.height(s, x, z)
        assert 32 <= h < 80
        col = g[z, :, x]
        assert np.all(col[h:] == 0)
        assert col[h - 1] == 1 and np.all(col[h - 4:h - 1] == 2) and np.all(col[:h - 4] == 3)
    assert set(np.unique(g)) <= {0, 1, 2, 3}


def test_blocks_and_sprites():
    g0 = S.generate(S.spec("c0"))
    t0 = S.generate(S.spec("c0", n_blocks=0))
    assert set(np.unique(g0[g0 != t0])) <= {4, 5, 6, 7, 8}
    # sprites: ~30% fill of 16^3 at the 70th percentile of each model's noise
    s = S.spec("c0", n_blocks=0, n_sprites=1, sprite_size=16)
    g = S.generate(s)
    t = S.generate(S.spec("c0", n_blocks=0))
    added = np.count_nonzero((g != t) & (g >= 4))
    assert 900 < added <= 1229


def test_oracle_outputs_pinned():
    g = S.generate(S.spec("c0"))
    cam, W, H = S.views("c0")[0]
    r = O.render(O.DenseWorld(g, S.palette()), cam, W, H, 512, 512, S.SUN)
    m = MANIFEST["oracle_c0"]
    for k in ("hits", "steps", "position", "normal", "illumination"):
        assert sha(r[k]) == m[k], k
    assert [int(x) for x in r["stats"]] == m["stats"]
    for p in m["pixels"]:
        assert [int(v) for v in r["hits"][p["y"], p["x"]]] == p["hit"]
        assert np.allclose(r["shaded"][p["y"], p["x"]], p["shaded"], atol=1e-6)
