"""CPU restatement of the reference's host-side world map -- TEST
INFRASTRUCTURE ONLY (the oracle for vt_world_set_voxels / get_voxels).

VoxelBrickmap (src/engine/voxel.zig:25-82) over GpuBlockAllocator
(src/engine/graphics/gpu_block_allocator.zig:4-41): a chunk table of
(dim/8)^3 u32 (0 = empty, else pool block index + 1, posToIndex order
x + M*(y + M*z)) and a pool of 512-word blocks handed out in allocation
order by get_block_for_chunk (:47-56); set (:58-61) writes word
x%8 + 8*(y%8 + 8*(z%8)) of the chunk's block, get (:63-70) reads it or 0.
"""
import numpy as np


class VoxelBrickmap:
    def __init__(self, dim: int):
        assert dim % 8 == 0
        self.dim, self.M = dim, dim // 8
        self.chunks = np.zeros(self.M ** 3, np.uint32)
        self.blocks = []  # one np.uint32[512] per allocated block

    def _chunk(self, x, y, z):
        return x // 8 + self.M * (y // 8 + self.M * (z // 8))

    def get_block_for_chunk(self, ci: int) -> int:
        if self.chunks[ci] > 0:
            return int(self.chunks[ci]) - 1
        self.blocks.append(np.zeros(512, np.uint32))
        self.chunks[ci] = len(self.blocks)
        return len(self.blocks) - 1

    def set(self, x, y, z, value):
        blk = self.get_block_for_chunk(self._chunk(x, y, z))
        self.blocks[blk][x % 8 + 8 * (y % 8 + 8 * (z % 8))] = value

    def get(self, x, y, z) -> int:
        idx = int(self.chunks[self._chunk(x, y, z)])
        return int(self.blocks[idx - 1][x % 8 + 8 * (y % 8 + 8 * (z % 8))]) if idx else 0

    def pool(self) -> np.ndarray:
        return (np.concatenate(self.blocks) if self.blocks else np.zeros(0, np.uint32))


def voxel(ty: int, solid: bool = True) -> int:
    """Voxel packed struct (voxel.zig:7-13): ty in bits 0-27, is_solid bit 28."""
    return (ty & 0xFFFFFFF) | (int(solid) << 28)
