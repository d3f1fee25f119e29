"""Model atlas from MagicaVoxel files (VoxelModelAtlas, src/engine/voxel.zig:84-131).

``Atlas`` holds the 256^3 RGBA8 texel array (host memory, index
x + 256*(y + 256*z)) that ``World.brickmap`` uploads; ``load`` places every
model of a .vox file into the next atlas slots through vt_atlas_load_vox.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

ATLAS_DIM = 256


def vox_info(data: bytes):
    n, ver = C.c_uint32(), C.c_uint32()
    N.check(N.lib().vt_vox_info(data, len(data), C.byref(n), C.byref(ver)), "vt_vox_info")
    return n.value, ver.value


class Atlas:
    def __init__(self):
        self.texels = np.zeros(ATLAS_DIM ** 3, dtype=np.uint32)
        self.next_slot = 0

    def load(self, data: bytes) -> range:
        """Places the file's models; returns their slot ids."""
        slot = C.c_uint32(self.next_slot)
        first = self.next_slot
        N.check(N.lib().vt_atlas_load_vox(data, len(data), self.texels.ctypes.data,
                                          C.byref(slot)), "vt_atlas_load_vox")
        self.next_slot = slot.value
        return range(first, self.next_slot)

    def load_file(self, path: str) -> range:
        with open(path, "rb") as f:
            return self.load(f.read())

    def slot_box(self, slot: int) -> np.ndarray:
        """(z, y, x)-indexed 8^3 view of one slot's texels."""
        a = self.texels.reshape(ATLAS_DIM, ATLAS_DIM, ATLAS_DIM)
        bx, by, bz = (slot % 32) * 8, (slot // 32 % 32) * 8, (slot // 1024 % 1024) * 8
        return a[bz:bz + 8, by:by + 8, bx:bx + 8]
