"""The engine's world layout (VoxelBrickmap + model atlas,
src/engine/voxel.zig:25-131) built from a dense material grid.

One grid voxel becomes one reference block whose value is its material k
(the Voxel ty field, voxel.zig:7-13) and atlas slot k holds a solid 8^3 cube
of palette[k], so traceMap over this layout returns the same hits as over
the dense world.  Chunks are numbered in increasing chunk-index order.
"""
from __future__ import annotations

import numpy as np

ATLAS_DIM = 256


def slot_origin(k: int):
    """Atlas slot origin (voxel.zig:100-102, map.glsl:58)."""
    return (k % 32) * 8, (k // 32 % 32) * 8, (k // 1024 % 32) * 8


def solid_atlas(palette: np.ndarray) -> np.ndarray:
    atlas = np.zeros((ATLAS_DIM,) * 3, np.uint32)
    for k in range(1, 256):
        ox, oy, oz = slot_origin(k)
        atlas[oz:oz + 8, oy:oy + 8, ox:ox + 8] = palette[k]
    return atlas.reshape(-1)


def from_dense(grid: np.ndarray, palette: np.ndarray):
    """(dim, chunks, pool, atlas) of a (dim, dim, dim)[z, y, x] uint8 grid."""
    dim = grid.shape[0]
    assert grid.shape == (dim, dim, dim) and dim % 8 == 0
    M = dim // 8
    flat = grid.reshape(-1)
    idx = np.flatnonzero(flat)
    x = idx % dim
    y = (idx // dim) % dim
    z = idx // (dim * dim)
    cid = (x >> 3) + M * ((y >> 3) + M * (z >> 3))
    used, rank = np.unique(cid, return_inverse=True)
    chunks = np.zeros(M ** 3, np.uint32)
    chunks[used] = np.arange(1, len(used) + 1, dtype=np.uint32)
    pool = np.zeros(len(used) * 512, np.uint32)
    pool[rank * 512 + (x & 7) + 8 * ((y & 7) + 8 * (z & 7))] = flat[idx]
    return dim, chunks, pool, solid_atlas(palette)
