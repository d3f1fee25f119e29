"""MagicaVoxel reader and model-atlas placement (vt_atlas_load_vox), CPU only.

Pins, without retyping the loader:
* a hand-built file with known voxels, palette and slots: the texel every
  voxel must land on is worked out by hand from voxel.zig:95-111 (slot
  origin, the (x, z, y) swizzle, colour index c -> palette entry c-1), plus
  the box clearing, the aliasing of coordinates >= 8 and the 512-entry cut;
* an independent chunk walk of the reference's own .vox assets (when the
  reference tree is present): model counts and voxels per model as listed in
  the asset survey, every placed texel a palette colour of its file, one
  texel per distinct in-box voxel index, and the 29 slots of the engine's
  load order (game.zig:101-113) -- digest and nonzero texels committed as
  tests/golden/atlas_assets.npz for the GPU tests.

zvox itself (zvox @ deb43ad) is not vendored: its byte order for colours
(file r, g, b, a -> R in the low byte) is the assumption that makes the
texel match packUnorm4x8(imageLoad) in map.glsl:57-60 -- parity unpinned
beyond that.
"""
import hashlib
import os
import struct

import numpy as np
import pytest

import vt
from vt import _native as N
from vt import vox

ASSETS = "/root/reference/assets"
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "atlas_assets.npz")
LOAD_ORDER = ["models", "grass", "grass2", "grass3", "grass4", "grass5", "rock", "flower",
              "water", "tree", "leaves", "dirt", "sand"]


def chunk(cid: bytes, content: bytes, children: bytes = b"") -> bytes:
    return cid + struct.pack("<II", len(content), len(children)) + content + children


def vox_file(models, palette=None, version=150) -> bytes:
    kids = b""
    for size, voxels in models:
        kids += chunk(b"SIZE", struct.pack("<III", *size))
        kids += chunk(b"XYZI", struct.pack("<I", len(voxels)) +
                      b"".join(bytes(v) for v in voxels))
    kids += chunk(b"nTRN", b"\x00" * 12)  # scene-graph chunks are skipped
    if palette is not None:
        kids += chunk(b"RGBA", b"".join(bytes(c) for c in palette))
    return b"VOX " + struct.pack("<I", version) + chunk(b"MAIN", b"", kids)


PAL = [(i, 255 - i, i // 2, 200) for i in range(256)]


def rgba(i):
    r, g, b, a = PAL[i]
    return r | (g << 8) | (b << 16) | (a << 24)


def at(x, y, z):
    return x + 256 * (y + 256 * z)


def test_known_answer_placement():
    data = vox_file([((8, 8, 8), [(1, 2, 3, 5), (0, 0, 0, 1), (7, 7, 7, 255)]),
                     ((8, 8, 8), [(2, 0, 1, 2), (9, 0, 0, 3), (0, 8, 0, 4)])], PAL)
    assert vox.vox_info(data) == (2, 150)
    a = vox.Atlas()
    a.texels[:] = 0xFFFFFFFF
    a.next_slot = 33
    assert list(a.load(data)) == [33, 34]
    assert a.next_slot == 35
    # slot 33: origin (8, 8, 0); voxel (x, y, z) -> (8 + x, 8 + z, 0 + y)
    assert a.texels[at(9, 11, 2)] == rgba(4)
    assert a.texels[at(8, 8, 0)] == rgba(0)
    assert a.texels[at(15, 15, 7)] == rgba(254)
    # slot 34: origin (16, 8, 0); (2, 0, 1) -> (18, 9, 0)
    assert a.texels[at(18, 9, 0)] == rgba(1)
    # (9, 0, 0): box index 9 aliases to box offset (1, 1, 0) -> (17, 9, 0)
    assert a.texels[at(17, 9, 0)] == rgba(2)
    # (0, 8, 0): box index 512 lies past the uploaded 8^3 box and is dropped
    box33, box34 = a.slot_box(33), a.slot_box(34)
    assert np.count_nonzero(box33) == 3 and np.count_nonzero(box34) == 2
    # the two boxes were cleared; every other texel is untouched
    rest = a.texels.reshape(256, 256, 256).copy()
    rest[0:8, 8:16, 8:24] = 0xFFFFFFFF
    assert np.all(rest == 0xFFFFFFFF)


def test_slot_origin_wraps_rows_and_layers():
    data = vox_file([((8, 8, 8), [(0, 0, 0, 1)])], PAL)
    for slot, origin in [(0, (0, 0, 0)), (31, (248, 0, 0)), (32, (0, 8, 0)), (1023, (248, 248, 0)),
                         (1024, (0, 0, 8)), (31 * 1024 + 5, (40, 0, 248))]:
        a = vox.Atlas()
        a.next_slot = slot
        a.load(data)
        assert a.texels[at(*origin)] == rgba(0), slot


def test_errors():
    a = vox.Atlas()
    with pytest.raises(vt.VtError) as e:
        a.load(vox_file([((8, 8, 8), [(0, 0, 0, 1)])], None))
    assert e.value.status == N.VT_ERR_UNSUPPORTED  # no RGBA chunk
    with pytest.raises(vt.VtError) as e:
        a.load(vox_file([((8, 8, 8), [(0, 0, 0, 0)])], PAL))
    assert e.value.status == N.VT_ERR_INVALID_ARGUMENT  # colour index 0
    good = vox_file([((8, 8, 8), [(0, 0, 0, 1)])], PAL)
    with pytest.raises(vt.VtError) as e:
        a.load(good[:-100])
    assert e.value.status == N.VT_ERR_INVALID_ARGUMENT  # truncated
    with pytest.raises(vt.VtError):
        a.load(b"NOPE" + good[4:])
    a.next_slot = 32 * 1024  # z origin 256: outside the atlas
    with pytest.raises(vt.VtError) as e:
        a.load(good)
    assert e.value.status == N.VT_ERR_INVALID_ARGUMENT and a.next_slot == 32 * 1024
    assert np.count_nonzero(a.texels) == 0  # nothing written by failed loads


# ---------------------------------------------------------------------------
# the reference's assets
# ---------------------------------------------------------------------------

def parse_vox(data: bytes):
    """Independent chunk walk: ([(size, voxels[n,4])], palette[256,4])."""
    assert data[:4] == b"VOX "
    models, pal, size = [], None, None

    def walk(off, end):
        nonlocal pal, size
        while off < end:
            cid = data[off:off + 4]
            n, m = struct.unpack_from("<II", data, off + 4)
            body = data[off + 12:off + 12 + n]
            if cid == b"MAIN":
                walk(off + 12 + n, off + 12 + n + m)
            elif cid == b"SIZE":
                size = struct.unpack_from("<III", body)
            elif cid == b"XYZI":
                cnt = struct.unpack_from("<I", body)[0]
                models.append((size, np.frombuffer(body[4:4 + 4 * cnt], np.uint8).reshape(-1, 4)))
            elif cid == b"RGBA":
                pal = np.frombuffer(body[:1024], np.uint8).reshape(256, 4)
            off += 12 + n + m

    walk(8, len(data))
    return models, pal


# asset survey: (models, min voxels, max voxels) per file
ASSET_COUNTS = {"models": (6, 385, 457), "grass": (1, 49, 49), "rock": (1, 176, 176),
          "flower": (1, 46, 46), "water": (1, 384, 384), "tree": (4, 229, 342),
          "leaves": (3, 331, 331), "dirt": (4, 443, 457), "sand": (4, 443, 457)}

needs_assets = pytest.mark.skipif(not os.path.isdir(ASSETS), reason="reference assets absent")


@needs_assets
def test_reference_assets_load_order_and_invariants():
    a = vox.Atlas()
    slots = {}
    for name in LOAD_ORDER:
        data = open(os.path.join(ASSETS, name + ".vox"), "rb").read()
        models, pal = parse_vox(data)
        assert vox.vox_info(data)[0] == len(models)
        if name in ASSET_COUNTS:
            n, lo, hi = ASSET_COUNTS[name]
            assert len(models) == n and all(lo <= len(v) <= hi for _, v in models), name
        slots[name] = a.load(data)
        colours = set((pal[:, 0].astype(np.uint32) | (pal[:, 1].astype(np.uint32) << 8) |
                       (pal[:, 2].astype(np.uint32) << 16) |
                       (pal[:, 3].astype(np.uint32) << 24)).tolist())
        for s, (size, v) in zip(slots[name], models):
            box = a.slot_box(s)
            idx = v[:, 0].astype(int) + 8 * (v[:, 2].astype(int) + 8 * v[:, 1].astype(int))
            assert np.count_nonzero(box) == len(set(idx[idx < 512].tolist())), (name, s)
            assert set(box[box != 0].tolist()) <= colours
    assert a.next_slot == 29
    assert list(slots["models"]) == list(range(0, 6)) and list(slots["sand"]) == [25, 26, 27, 28]
    assert list(slots["tree"]) == [14, 15, 16, 17] and list(slots["leaves"]) == [18, 19, 20]
    nz = np.flatnonzero(a.texels).astype(np.uint32)
    g = np.load(GOLDEN)
    assert np.array_equal(g["index"], nz) and np.array_equal(g["value"], a.texels[nz])
    assert hashlib.sha256(a.texels.tobytes()).hexdigest() == str(g["sha256"])


def test_golden_atlas_fixture_is_consistent():
    g = np.load(GOLDEN)
    idx, val = g["index"], g["value"]
    assert idx.dtype == np.uint32 and np.all(np.diff(idx.astype(np.int64)) > 0)
    t = np.zeros(256 ** 3, np.uint32)
    t[idx] = val
    assert hashlib.sha256(t.tobytes()).hexdigest() == str(g["sha256"])
    # 29 slots of 8^3, all in the first layer of 32x32 slots
    z = idx // 65536
    assert z.max() < 8 and np.all(val != 0)
