// vt_vox.hip -- MagicaVoxel .vox reader and model-atlas placement (host code).
//
// Replaces VoxelModelAtlas.load_block_model / load_single_block_model
// (src/engine/voxel.zig:94-127) over the zvox package (zvox @ deb43ad,
// build.zig.zon:27-30, not vendored).  The file format is restated from the
// published MagicaVoxel format: "VOX " + version, then chunks of
// {id[4], content bytes N, children bytes M}; MAIN's children hold SIZE /
// XYZI pairs (one per model, in file order) and one RGBA palette chunk of
// 256 colours whose entry i is colour index i + 1; every other chunk (PACK,
// nTRN, nGRP, nSHP, LAYR, MATL, rOBJ, rCAM, NOTE, IMAP) is skipped.
//
// Placement follows voxel.zig:95-111: model k of the atlas (k = the running
// slot counter) owns the 8^3 box at ((k % 32) * 8, ((k / 32) % 32) * 8,
// ((k / 1024) % 1024) * 8); the box is cleared, then voxel (x, y, z, c) lands
// at box index x + 8 * (z + 8 * y) -- atlas (bx + x, by + z, bz + y), the
// .vox z axis becomes the atlas y axis -- with colour palette[c - 1] as an
// RGBA8 texel (the file's r, g, b, a bytes, R in the low byte).  Like the
// reference, coordinates are not range-checked against 8: a box index >= 512
// is dropped (only the 8^3 box is uploaded) and later voxels overwrite
// earlier ones.
#include <cstring>

#include "vt_internal.hpp"

namespace {

struct Reader {
  const uint8_t* p;
  size_t n, at;
  bool u32(uint32_t& v) {
    if (n - at < 4) return false;
    v = (uint32_t)p[at] | ((uint32_t)p[at + 1] << 8) | ((uint32_t)p[at + 2] << 16) |
        ((uint32_t)p[at + 3] << 24);
    at += 4;
    return true;
  }
};

struct Chunk {
  char id[5];
  size_t content, content_len, children, children_len, end;
};

bool read_chunk(const uint8_t* p, size_t len, size_t at, Chunk& c) {
  if (len - at < 12 || at > len) return false;
  std::memcpy(c.id, p + at, 4);
  c.id[4] = 0;
  Reader r{p, len, at + 4};
  uint32_t n, m;
  if (!r.u32(n) || !r.u32(m)) return false;
  c.content = at + 12;
  c.content_len = n;
  if (len - c.content < n) return false;
  c.children = c.content + n;
  c.children_len = m;
  if (len - c.children < m) return false;
  c.end = c.children + m;
  return true;
}

struct Model {
  uint32_t sx, sy, sz;
  size_t xyzi;  // offset of the voxel records
  uint32_t count;
};

}  // namespace

extern "C" int vt_vox_info(const uint8_t* bytes, size_t len, uint32_t* n_models,
                           uint32_t* version) {
  using vt::fail;
  if (!bytes || !n_models) return fail(VT_ERR_NULL_POINTER, "vt_vox_info: null argument");
  if (len < 8 || std::memcmp(bytes, "VOX ", 4) != 0)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_vox_info: not a .vox file");
  Reader r{bytes, len, 4};
  uint32_t ver = 0;
  r.u32(ver);
  Chunk main;
  if (!read_chunk(bytes, len, 8, main) || std::strcmp(main.id, "MAIN") != 0)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_vox_info: no MAIN chunk");
  uint32_t models = 0;
  for (size_t at = main.children; at < main.end;) {
    Chunk c;
    if (!read_chunk(bytes, main.end, at, c))
      return fail(VT_ERR_INVALID_ARGUMENT, "vt_vox_info: truncated chunk at %zu", at);
    if (std::strcmp(c.id, "XYZI") == 0) models++;
    at = c.end;
  }
  *n_models = models;
  if (version) *version = ver;
  return VT_OK;
}

extern "C" int vt_atlas_load_vox(const uint8_t* bytes, size_t len, uint32_t* atlas_rgba8,
                                 uint32_t* next_slot) {
  using vt::fail;
  if (!bytes || !atlas_rgba8 || !next_slot)
    return fail(VT_ERR_NULL_POINTER, "vt_atlas_load_vox: null argument");
  if (len < 8 || std::memcmp(bytes, "VOX ", 4) != 0)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: not a .vox file");
  Chunk main;
  if (!read_chunk(bytes, len, 8, main) || std::strcmp(main.id, "MAIN") != 0)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: no MAIN chunk");
  // pass 1: models (SIZE then XYZI) and the palette
  Model models[1024];
  uint32_t nm = 0;
  const uint8_t* rgba = nullptr;
  bool have_size = false;
  Model cur{};
  for (size_t at = main.children; at < main.end;) {
    Chunk c;
    if (!read_chunk(bytes, main.end, at, c))
      return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: truncated chunk at %zu", at);
    Reader r{bytes, c.content + c.content_len, c.content};
    if (std::strcmp(c.id, "SIZE") == 0) {
      if (!r.u32(cur.sx) || !r.u32(cur.sy) || !r.u32(cur.sz))
        return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: short SIZE chunk");
      have_size = true;
    } else if (std::strcmp(c.id, "XYZI") == 0) {
      if (!have_size) return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: XYZI before SIZE");
      if (!r.u32(cur.count) || (c.content_len - 4) / 4 < cur.count)
        return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: short XYZI chunk");
      cur.xyzi = c.content + 4;
      if (nm == 1024) return fail(VT_ERR_UNSUPPORTED, "vt_atlas_load_vox: over 1024 models");
      models[nm++] = cur;
      have_size = false;
    } else if (std::strcmp(c.id, "RGBA") == 0) {
      if (c.content_len < 1024)
        return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: short RGBA chunk");
      rgba = bytes + c.content;
    }
    at = c.end;
  }
  if (!rgba)
    return fail(VT_ERR_UNSUPPORTED,
                "vt_atlas_load_vox: no RGBA chunk (the default palette is not supported)");
  // every model's box must fit the 256^3 atlas before anything is written
  const uint32_t first = *next_slot;
  for (uint32_t k = 0; k < nm; k++) {
    const uint32_t slot = first + k;
    if (((slot / 1024u) % 1024u) * 8u + 8u > 256u)
      return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: slot %u is outside the atlas",
                  slot);
    const uint8_t* v = bytes + models[k].xyzi;
    for (uint32_t i = 0; i < models[k].count; i++)
      if (v[4 * i + 3] == 0)
        return fail(VT_ERR_INVALID_ARGUMENT, "vt_atlas_load_vox: colour index 0 in model %u",
                    k);
  }
  // pass 2: place (voxel.zig:95-111)
  for (uint32_t k = 0; k < nm; k++) {
    const uint32_t slot = first + k;
    const uint32_t bx = (slot % 32u) * 8u, by = ((slot / 32u) % 32u) * 8u,
                   bz = ((slot / 1024u) % 1024u) * 8u;
    uint32_t box[512];
    std::memset(box, 0, sizeof box);
    const uint8_t* v = bytes + models[k].xyzi;
    for (uint32_t i = 0; i < models[k].count; i++) {
      const uint32_t x = v[4 * i], y = v[4 * i + 1], z = v[4 * i + 2], ci = v[4 * i + 3];
      const uint32_t idx = x + 8u * (z + 8u * y);
      if (idx >= 512u) continue;
      const uint8_t* col = rgba + 4 * (ci - 1u);
      box[idx] = (uint32_t)col[0] | ((uint32_t)col[1] << 8) | ((uint32_t)col[2] << 16) |
                 ((uint32_t)col[3] << 24);
    }
    for (uint32_t k2 = 0; k2 < 8; k2++)
      for (uint32_t j = 0; j < 8; j++)
        for (uint32_t i = 0; i < 8; i++)
          atlas_rgba8[(size_t)(bx + i) + 256u * ((size_t)(by + j) + 256u * (bz + k2))] =
              box[i + 8u * (j + 8u * k2)];
  }
  *next_slot = first + nm;
  return VT_OK;
}
