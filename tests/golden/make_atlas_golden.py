"""Writes tests/golden/atlas_assets.npz: the model atlas after loading the
reference's .vox assets in the engine's order (game.zig:101-113) through
vt_atlas_load_vox -- nonzero texel indices and values plus the sha256 of the
whole 256^3 texel array.  Run in a tree that has the reference's assets; the
data files themselves are not copied, only the placed texels."""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "unnamed-voxel-tracer_amd"))
from vt import vox  # noqa: E402

ASSETS = "/root/reference/assets"
ORDER = ["models", "grass", "grass2", "grass3", "grass4", "grass5", "rock", "flower", "water",
         "tree", "leaves", "dirt", "sand"]

a = vox.Atlas()
for name in ORDER:
    a.load_file(os.path.join(ASSETS, name + ".vox"))
nz = np.flatnonzero(a.texels).astype(np.uint32)
np.savez_compressed(os.path.join(HERE, "atlas_assets.npz"), index=nz, value=a.texels[nz],
                    sha256=np.array(hashlib.sha256(a.texels.tobytes()).hexdigest()),
                    slots=np.array(a.next_slot))
print(len(nz), "texels in", a.next_slot, "slots")
