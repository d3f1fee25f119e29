"""Builders of the reference world layout (chunk table + block pool + 256^3
RGBA8 model atlas, src/engine/voxel.zig:25-131) from dense grids, for tests.

Embedding A: one grid voxel = one reference block; block value = material k;
  atlas slot k = a solid 8^3 cube of palette[k].  traceMap never descends.
Embedding C: one grid voxel = one sub-voxel; each nonempty 8^3 brick of the
  grid becomes one atlas model (deduplicated), MAP_DIMENSION = dim / 8, so
  traceMap runs both levels.
"""
import numpy as np


def slot_origin(k):
    # voxel.zig:100-102 / map.glsl:58
    return (k % 32) * 8, (k // 32 % 32) * 8, (k // 1024 % 32) * 8


def chunked(blocks):
    """chunk table + pool from a (n,n,n)[z,y,x] uint32 block grid, chunks
    allocated in first-touch order of x-fastest iteration (VoxelBrickmap.set)."""
    n = blocks.shape[0]
    M = n // 8
    chunks = np.zeros(M ** 3, np.uint32)
    pool = []
    zs, ys, xs = np.nonzero(blocks)
    order = np.lexsort((xs, ys, zs))
    for i in order:
        x, y, z = xs[i], ys[i], zs[i]
        ci = x // 8 + M * (y // 8 + (z // 8) * M)
        if chunks[ci] == 0:
            pool.append(np.zeros(512, np.uint32))
            chunks[ci] = len(pool)
        pool[chunks[ci] - 1][x % 8 + ((y % 8) + (z % 8) * 8) * 8] = blocks[z, y, x]
    pool = np.concatenate(pool) if pool else np.zeros(0, np.uint32)
    return chunks, pool


def embedding_a(grid, palette):
    chunks, pool = chunked(grid.astype(np.uint32))
    atlas = np.zeros((256, 256, 256), np.uint32)
    for k in range(1, 256):
        ox, oy, oz = slot_origin(k)
        atlas[oz:oz + 8, oy:oy + 8, ox:ox + 8] = palette[k]
    return grid.shape[0], chunks, pool, atlas.ravel()


def embedding_c(grid, palette):
    n = grid.shape[0]
    nb = n // 8
    atlas = np.zeros((256, 256, 256), np.uint32)
    blocks = np.zeros((nb, nb, nb), np.uint32)
    models = {}
    for bz in range(nb):
        for by in range(nb):
            for bx in range(nb):
                sub = grid[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8]
                if not sub.any():
                    continue
                key = sub.tobytes()
                if key not in models:
                    k = len(models)
                    assert k < 32768
                    models[key] = k
                    ox, oy, oz = slot_origin(k)
                    atlas[oz:oz + 8, oy:oy + 8, ox:ox + 8] = np.where(sub != 0, palette[sub], 0)
                # Voxel packing (voxel.zig:7-13): ty in the low 28 bits, is_solid bit 28
                blocks[bz, by, bx] = models[key] | (1 << 28)
    chunks, pool = chunked(blocks)
    return nb, chunks, pool, atlas.ravel()
