// vt_api.hip -- the extern "C" entry points declared in include/vt.h.
//
// Validation happens here, on the host, before anything is enqueued: every
// size, pointer and alignment a kernel relies on is checked so that no launch
// can index outside the buffers it is given.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "vt_internal.hpp"

namespace {
thread_local char g_detail[512] = "";
}  // namespace

namespace vt {

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_detail, sizeof g_detail, fmt, ap);
  va_end(ap);
  return status;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(VT_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

bool aligned(const void* p, size_t a) { return ((uintptr_t)p & (a - 1)) == 0; }

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace vt

namespace {
using vt::aligned;
using vt::dev_free;
using vt::fail;
using vt::hip_fail;

void free_world(vt_world* w) {
  if (!w) return;
  dev_free(w->prefix);
  dev_free(w->vox);
  dev_free(w->l0);
  dev_free(w->l1);
  dev_free(w->l1b);
  dev_free(w->palette);
  dev_free(w->chunks);
  dev_free(w->pool);
  dev_free(w->atlas);
  dev_free(w->occ);
  dev_free(w->brick_bits);
  dev_free(w->freemap);
  dev_free(w->free_tmp);
  delete[] w->chunk_mirror;
  delete w;
}

vt_world* new_world(vt_ctx* ctx, WorldKind kind, uint32_t dim) {
  vt_world* w = new (std::nothrow) vt_world;
  if (!w) return nullptr;
  std::memset(w, 0, sizeof *w);
  w->ctx = ctx;
  w->kind = kind;
  w->dim = dim;
  return w;
}

int dev_alloc(vt_world* w, void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(VT_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) for %s: %s", bytes, what,
                hipGetErrorString(e));
  }
  w->bytes += bytes;
  return VT_OK;
}

// The free-region map and its build scratch (built by vt::rebuild_freemap).
int alloc_freemap(vt_world* w) {
  const uint32_t R = vt::free_regions(w->dim);
  int st;
  if ((st = dev_alloc(w, (void**)&w->freemap, 8u * vt::free_words_per_octant(w->dim) * 4u,
                      "free-region map")) != VT_OK)
    return st;
  return dev_alloc(w, (void**)&w->free_tmp, (size_t)8u * R * R * R, "free-region scratch");
}

}  // namespace

namespace vt {

DenseWorld dense_view(const vt_world* w) {
  DenseWorld d;
  d.vox = w->vox;
  d.l0 = w->l0;
  d.l1 = w->l1;
  d.l1b = w->l1b;
  d.palette = w->palette;
  d.dim = w->dim;
  uint32_t bits = 0;
  while ((1u << bits) < w->dim) bits++;
  const uint32_t low = bits * 3 >= 32 ? 0xFFFFFFFFu : ((1u << (bits * 3)) - 1u);
  d.log2dim = bits;
  d.mx = 0x09249249u & low;
  d.my = (0x09249249u << 1) & low;
  d.mz = (0x09249249u << 2) & low;
  d.nvox = w->dim * w->dim * w->dim;
  d.prefix = w->prefix;
  d.freemap = w->freemap;
  d.free_oct_bytes = free_words_per_octant(w->dim) * 4u;
  d.free_bytes = w->freemap && w->free_valid ? 8u * d.free_oct_bytes : 0u;
  reach_of(w, d.reach_lo, d.reach_hi);
  return d;
}

// The box of the nonempty voxels widened by wide_margin (the whole grid
// while the world has none: no ray ends early for the box).
void reach_of(const vt_world* w, float* lo, float* hi) {
  const float m = wide_margin(w->dim);
  for (int i = 0; i < 3; i++) {
    lo[i] = w->box_any ? (float)w->box_lo[i] - m : 0.0f;
    hi[i] = w->box_any ? (float)w->box_hi[i] + 1.0f + m : (float)w->dim;
  }
}

void grow_box(vt_world* w, int32_t x, int32_t y, int32_t z) {
  const int32_t c[3] = {x, y, z};
  for (int a = 0; a < 3; a++) {
    w->box_lo[a] = w->box_any ? std::min(w->box_lo[a], c[a]) : c[a];
    w->box_hi[a] = w->box_any ? std::max(w->box_hi[a], c[a]) : c[a];
  }
  w->box_any = true;
}

// The box of the nonzero blocks of a reference-layout map: exact from the
// pool when `pool` is given, else the box of the nonzero chunk-table entries
// (8^3 blocks each).
void scan_brickmap_box(const uint32_t* chunks, const uint32_t* pool, uint32_t dim, vt_world* w) {
  w->box_any = false;
  const uint32_t M = dim / 8u;
  for (uint32_t cz = 0; cz < M; cz++)
    for (uint32_t cy = 0; cy < M; cy++)
      for (uint32_t cx = 0; cx < M; cx++) {
        const uint32_t e = chunks[cx + M * (cy + M * cz)];
        if (!e) continue;
        if (!pool) {
          grow_box(w, (int32_t)(8u * cx), (int32_t)(8u * cy), (int32_t)(8u * cz));
          grow_box(w, (int32_t)(8u * cx + 7u), (int32_t)(8u * cy + 7u), (int32_t)(8u * cz + 7u));
          continue;
        }
        const uint32_t* b = pool + (size_t)(e - 1u) * 512u;  // in-block x + 8y + 64z
        for (uint32_t i = 0; i < 512u; i++)
          if (b[i])
            grow_box(w, (int32_t)(8u * cx + (i & 7u)), (int32_t)(8u * cy + ((i >> 3) & 7u)),
                     (int32_t)(8u * cz + (i >> 6)));
      }
}

// The box of the nonempty voxels of a linear grid (x + dim (y + dim z)).
void scan_dense_box(const uint8_t* grid, uint32_t dim, vt_world* w) {
  w->box_any = false;
  for (int i = 0; i < 3; i++) {
    w->box_lo[i] = (int32_t)dim;
    w->box_hi[i] = -1;
  }
  for (uint32_t z = 0; z < dim; z++)
    for (uint32_t y = 0; y < dim; y++) {
      const uint8_t* row = grid + (size_t)dim * (y + (size_t)dim * z);
      uint64_t any = 0;
      for (uint32_t x = 0; x + 8 <= dim; x += 8) {
        uint64_t q;
        std::memcpy(&q, row + x, 8);
        any |= q;
      }
      for (uint32_t x = dim & ~7u; x < dim; x++) any |= row[x];
      if (!any) continue;
      uint32_t x0 = 0, x1 = dim - 1;
      while (!row[x0]) x0++;
      while (!row[x1]) x1--;
      w->box_any = true;
      w->box_lo[0] = std::min(w->box_lo[0], (int32_t)x0);
      w->box_hi[0] = std::max(w->box_hi[0], (int32_t)x1);
      w->box_lo[1] = std::min(w->box_lo[1], (int32_t)y);
      w->box_hi[1] = std::max(w->box_hi[1], (int32_t)y);
      w->box_lo[2] = std::min(w->box_lo[2], (int32_t)z);
      w->box_hi[2] = std::max(w->box_hi[2], (int32_t)z);
    }
}

BrickmapWorld brickmap_view(const vt_world* w) {
  BrickmapWorld b;
  b.chunks = w->chunks;
  b.pool = w->pool;
  b.atlas = w->atlas;
  b.dim = w->dim;
  b.occ = w->occ;
  b.nvox = w->dim * w->dim * w->dim;
  uint32_t bits = 0;
  while ((1u << bits) < w->dim) bits++;
  const uint32_t low = bits * 3 >= 32 ? 0xFFFFFFFFu : ((1u << (bits * 3)) - 1u);
  b.mx = 0x09249249u & low;
  b.my = (0x09249249u << 1) & low;
  b.mz = (0x09249249u << 2) & low;
  b.freemap = w->freemap;
  b.free_oct_bytes = free_words_per_octant(w->dim) * 4u;
  b.free_bytes = w->freemap && w->free_valid ? 8u * b.free_oct_bytes : 0u;
  reach_of(w, b.reach_lo, b.reach_hi);
  return b;
}

}  // namespace vt

namespace vt {

int check_image_size(uint32_t W, uint32_t H, const char* who) {
  if (W == 0 || H == 0 || W > 16384 || H > 16384)
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: image %ux%u outside [1,16384]^2", who, W, H);
  return VT_OK;
}

int check_params(const vt_render_params& p, const vt_world* world, const char* who) {
  if (p.max_steps_primary < 0 || p.max_steps_primary > 65535 || p.max_steps_shadow < 0 ||
      p.max_steps_shadow > 65535)
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: step caps must lie in [0, 65535]", who);
  if ((p.flags & VT_FLAG_FREE_MAP_ON) && (p.flags & VT_FLAG_FREE_MAP_OFF))
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: FREE_MAP_ON and FREE_MAP_OFF together", who);
  if (p.flags & ~(VT_FLAG_FREE_MAP_ON | VT_FLAG_FREE_MAP_OFF | VT_FLAG_GBUF_UNORM8 |
                  VT_FLAG_LAMBERT | VT_FLAG_ALBEDO_COMPACT |
                  VT_FLAG_GRAD_L2 | VT_FLAG_PRESENT))
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: unknown flags 0x%x", who, p.flags);
  if ((p.flags & VT_FLAG_PRESENT) && (p.flags & VT_FLAG_LAMBERT))
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: VT_FLAG_PRESENT applies to the composite, not "
                "to VT_FLAG_LAMBERT", who);
  if ((p.flags & (VT_FLAG_ALBEDO_COMPACT | VT_FLAG_GRAD_L2)) && !(p.flags & VT_FLAG_LAMBERT))
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: ALBEDO_COMPACT / GRAD_L2 need VT_FLAG_LAMBERT", who);
  if ((p.flags & VT_FLAG_ALBEDO_COMPACT) && (world->kind != WorldKind::Dense || !world->prefix))
    return fail(VT_ERR_UNSUPPORTED,
                "%s: VT_FLAG_ALBEDO_COMPACT needs vt_world_build_compact_index first", who);
  if ((p.flags & VT_FLAG_GRAD_L2) && !p.grad_albedo)
    return fail(VT_ERR_NULL_POINTER, "%s: VT_FLAG_GRAD_L2 needs grad_albedo", who);
  if ((p.grad_albedo && !aligned(p.grad_albedo, 4)) || (p.grad_sun && !aligned(p.grad_sun, 4)) ||
      (p.loss && !aligned(p.loss, 4)))
    return fail(VT_ERR_MISALIGNED, "%s: gradient outputs", who);
  if (p.traversal > VT_TRAVERSAL_BITMASK8)
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: unknown traversal %u", who, p.traversal);
  if (world->kind == WorldKind::Brickmap && p.traversal != VT_TRAVERSAL_FLAT)
    return fail(VT_ERR_UNSUPPORTED, "%s: brickmap worlds trace with VT_TRAVERSAL_FLAT", who);
  if ((p.flags & VT_FLAG_LAMBERT) && (!p.albedo || world->kind != WorldKind::Dense))
    return fail(VT_ERR_UNSUPPORTED, "%s: VT_FLAG_LAMBERT needs a dense world and albedo", who);
  if (p.albedo && !aligned(p.albedo, 4)) return fail(VT_ERR_MISALIGNED, "%s: albedo", who);
  for (int i = 0; i < 3; i++)
    if (!std::isfinite(p.sun[i])) return fail(VT_ERR_INVALID_ARGUMENT, "%s: sun not finite", who);
  if ((p.flags & VT_FLAG_LAMBERT) && !sun_normalizable(p.sun))
    return fail(VT_ERR_INVALID_ARGUMENT,
                "%s: VT_FLAG_LAMBERT normalises the sun: its squared length must be a nonzero "
                "finite binary32 value", who);
  return VT_OK;
}

bool sun_normalizable(const float sun[3]) {
  // the kernels' len3: x*x + y*y + z*z rounded per operation
  float s = sun[0] * sun[0];
  s = s + sun[1] * sun[1];
  s = s + sun[2] * sun[2];
  return std::isfinite(s) && s > 0.0f;
}

int check_frame(const vt_frame* frame, uint32_t flags, const char* who) {
  if ((flags & VT_FLAG_GRAD_L2) && !frame->target)
    return fail(VT_ERR_NULL_POINTER, "%s: VT_FLAG_GRAD_L2 needs frame->target", who);
  if (frame->display && !(flags & VT_FLAG_PRESENT))
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: frame->display needs VT_FLAG_PRESENT", who);
  if (frame->display && !frame->shaded)
    return fail(VT_ERR_NULL_POINTER, "%s: frame->display is written with frame->shaded", who);
  const void* f16[] = {frame->shaded, frame->color, frame->normal, frame->position,
                       frame->illumination, frame->target};
  for (const void* q : f16)
    if (q && !aligned(q, 16)) return fail(VT_ERR_MISALIGNED, "%s: image not 16B aligned", who);
  if ((frame->hits && !aligned(frame->hits, 4)) || (frame->steps && !aligned(frame->steps, 4)) ||
      (frame->display && !aligned(frame->display, 4)) ||
      (frame->schedule && !aligned(frame->schedule, 4)) ||
      (frame->stats && !aligned(frame->stats, 8)))
    return fail(VT_ERR_MISALIGNED, "%s: hits/steps/display/schedule/stats misaligned", who);
  return VT_OK;
}

ViewConsts view_consts(const vt_camera* camera, uint32_t W, uint32_t H) {
  ViewConsts v;
  std::memcpy(v.pos, camera->position, sizeof v.pos);
  std::memcpy(v.view, camera->view, sizeof v.view);
  v.tan_half = (float)std::tan((double)(camera->fov / 2.0f));
  v.W = W;
  v.H = H;
  return v;
}

ShadeConsts shade_consts(const vt_render_params& p) {
  ShadeConsts sc;
  std::memcpy(sc.sun, p.sun, sizeof sc.sun);
  // the dense launchers redo this with the world's axis masks
  sc.shadow = make_ray_consts(p.sun[0], p.sun[1], p.sun[2], 0u, 0u, 0u);
  sc.max_primary = p.max_steps_primary;
  sc.max_shadow = p.max_steps_shadow;
  sc.flags = p.flags;
  sc.ambient = p.ambient;
  sc.albedo = p.albedo;
  sc.grad_albedo = p.grad_albedo;
  sc.grad_sun = p.grad_sun;
  sc.loss = p.loss;
  return sc;
}

bool render_free_map(const vt_world* w, const vt_render_params& p, const vt_frame& f) {
  if (f.steps) return false;  // step counts are stored
  if (w->kind == WorldKind::Dense) {
    if (p.traversal != VT_TRAVERSAL_FLAT) return false;
    // statistics count traceMap's steps unless the test is forced on (then
    // the steps executed)
    if (f.stats && !(p.flags & VT_FLAG_FREE_MAP_ON)) return false;
  } else if (f.stats || !w->occ) {  // the reference layout's statistics kernel is the literal loop
    return false;
  }
  return use_free_map(w, shade_consts(p));
}

FrameOut frame_out(const vt_frame* frame) {
  FrameOut o;
  o.shaded = reinterpret_cast<float4*>(frame->shaded);
  o.hits = reinterpret_cast<uint32_t*>(frame->hits);
  o.color = reinterpret_cast<float4*>(frame->color);
  o.normal = reinterpret_cast<float4*>(frame->normal);
  o.position = reinterpret_cast<float4*>(frame->position);
  o.illum = reinterpret_cast<float4*>(frame->illumination);
  o.steps = frame->steps;
  o.stats = frame->stats;
  o.target = reinterpret_cast<const float4*>(frame->target);
  o.display = frame->display;
  return o;
}

}  // namespace vt

extern "C" {

const char* vt_status_string(int status) {
  switch (status) {
    case VT_OK: return "ok";
    case VT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case VT_ERR_NULL_POINTER: return "null pointer";
    case VT_ERR_OUT_OF_MEMORY: return "out of memory";
    case VT_ERR_DEVICE: return "HIP runtime error";
    case VT_ERR_NO_DEVICE: return "no HIP device";
    case VT_ERR_UNSUPPORTED: return "unsupported for this world or mode";
    case VT_ERR_MISALIGNED: return "misaligned device pointer";
    case VT_ERR_WRONG_CONTEXT: return "world belongs to another context";
    default: return "unknown status";
  }
}

const char* vt_last_error(void) { return g_detail; }

int vt_create(vt_ctx** out, int device) {
  if (!out) return fail(VT_ERR_NULL_POINTER, "vt_create: out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(VT_ERR_NO_DEVICE, "vt_create: no HIP device (%s)", hipGetErrorString(e));
  }
  if (device < 0 || device >= n)
    return fail(VT_ERR_NO_DEVICE, "vt_create: device %d of %d", device, n);
  VT_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  VT_HIP(hipGetDeviceProperties(&prop, device));
  vt_ctx* c = new (std::nothrow) vt_ctx;
  if (!c) return fail(VT_ERR_OUT_OF_MEMORY, "vt_create: host allocation");
  c->device = device;
  c->compute_units = prop.multiProcessorCount;
  std::snprintf(c->name, sizeof c->name, "%s", prop.gcnArchName);
  c->stage = nullptr;
  e = hipMalloc(&c->stage, kCtxStageBytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return fail(VT_ERR_OUT_OF_MEMORY, "vt_create: %zu bytes of staging: %s", kCtxStageBytes,
                hipGetErrorString(e));
  }
  int st = vt::noalloc_init(device);
  if (st != VT_OK) {
    dev_free(c->stage);
    delete c;
    return st;
  }
  *out = c;
  return VT_OK;
}

int vt_destroy(vt_ctx* ctx) {
  if (!ctx) return fail(VT_ERR_NULL_POINTER, "vt_destroy: ctx is NULL");
  (void)hipSetDevice(ctx->device);
  dev_free(ctx->stage);
  delete ctx;
  return VT_OK;
}

int vt_sync(vt_ctx* ctx, void* stream) {
  if (!ctx) return fail(VT_ERR_NULL_POINTER, "vt_sync: ctx is NULL");
  VT_HIP(hipSetDevice(ctx->device));
  VT_HIP(hipStreamSynchronize((hipStream_t)stream));
  return VT_OK;
}

int vt_device_info(const vt_ctx* ctx, int* compute_units, char* name, size_t name_len) {
  if (!ctx) return fail(VT_ERR_NULL_POINTER, "vt_device_info: ctx is NULL");
  if (compute_units) *compute_units = ctx->compute_units;
  if (name && name_len) std::snprintf(name, name_len, "%s", ctx->name);
  return VT_OK;
}

int vt_world_create_dense_u8(vt_ctx* ctx, const uint8_t* grid, uint32_t dim,
                             const uint32_t palette_rgba8[256], vt_world** out) {
  if (!out) return fail(VT_ERR_NULL_POINTER, "vt_world_create_dense_u8: out is NULL");
  *out = nullptr;
  if (!ctx || !grid || !palette_rgba8)
    return fail(VT_ERR_NULL_POINTER, "vt_world_create_dense_u8: ctx, grid and palette required");
  if (dim < 16 || dim > 1024 || (dim & (dim - 1)) != 0)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_create_dense_u8: dim %u must be a power of two in [16, 1024]", dim);
  for (int k = 1; k < 256; k++)
    if (palette_rgba8[k] == 0u)
      return fail(VT_ERR_INVALID_ARGUMENT,
                  "vt_world_create_dense_u8: palette[%d] is 0 (a material colour must be "
                  "nonzero: a zero texel reads as empty)",
                  k);
  VT_HIP(hipSetDevice(ctx->device));
  vt_world* w = new_world(ctx, WorldKind::Dense, dim);
  if (!w) return fail(VT_ERR_OUT_OF_MEMORY, "vt_world_create_dense_u8: host allocation");
  const size_t n = (size_t)dim * dim * dim;
  int st;
  if ((st = dev_alloc(w, (void**)&w->vox, n, "voxels")) != VT_OK ||
      (st = dev_alloc(w, (void**)&w->l0, n / 8, "L0 masks")) != VT_OK ||
      (st = dev_alloc(w, (void**)&w->l1, n / 512, "L1 masks")) != VT_OK ||
      (st = dev_alloc(w, (void**)&w->l1b, vt::l1b_words(dim) * 8, "8^3-brick masks")) != VT_OK ||
      (st = dev_alloc(w, (void**)&w->palette, 256 * 4, "palette")) != VT_OK ||
      (st = alloc_freemap(w)) != VT_OK) {
    free_world(w);
    return st;
  }
  uint8_t* lin = nullptr;
  hipError_t e = hipMalloc((void**)&lin, n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    free_world(w);
    return fail(VT_ERR_OUT_OF_MEMORY, "vt_world_create_dense_u8: staging %zu bytes", n);
  }
  vt::scan_dense_box(grid, dim, w);
  e = hipMemcpy(lin, grid, n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(w->palette, palette_rgba8, 256 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = vt::launch_build_dense(lin, dim, w->vox, w->l0, w->l1, w->l1b, nullptr);
  if (e == hipSuccess) e = vt::rebuild_freemap(w, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e == hipSuccess) e = vt::finish_freemap(w);
  dev_free(lin);
  if (e != hipSuccess) {
    free_world(w);
    return hip_fail(e, "vt_world_create_dense_u8: upload/build");
  }
  *out = w;
  return VT_OK;
}

int vt_world_create_brickmap(vt_ctx* ctx, uint32_t dim, const uint32_t* chunks,
                             const uint32_t* pool, size_t pool_words,
                             const uint32_t* atlas_rgba8, vt_world** out) {
  if (!out) return fail(VT_ERR_NULL_POINTER, "vt_world_create_brickmap: out is NULL");
  *out = nullptr;
  if (!ctx || !chunks || !atlas_rgba8 || (!pool && pool_words))
    return fail(VT_ERR_NULL_POINTER, "vt_world_create_brickmap: ctx, chunks, atlas required");
  if (dim < 8 || dim > 1024 || dim % 8 != 0)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_create_brickmap: dim %u must be a multiple of 8 in [8, 1024]", dim);
  if (pool_words % 512 != 0)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_create_brickmap: pool_words %zu is not a whole number of 8^3 blocks",
                pool_words);
  const size_t M = dim / 8, nchunks = M * M * M, nblocks = pool_words / 512;
  for (size_t i = 0; i < nchunks; i++)
    if (chunks[i] > nblocks)
      return fail(VT_ERR_INVALID_ARGUMENT,
                  "vt_world_create_brickmap: chunk %zu names pool block %u of %zu", i,
                  chunks[i] - 1u, nblocks);
  VT_HIP(hipSetDevice(ctx->device));
  vt_world* w = new_world(ctx, WorldKind::Brickmap, dim);
  if (!w) return fail(VT_ERR_OUT_OF_MEMORY, "vt_world_create_brickmap: host allocation");
  const size_t atlas_words = 256u * 256u * 256u;
  int st;
  if ((st = dev_alloc(w, (void**)&w->chunks, nchunks * 4, "chunk table")) != VT_OK ||
      (st = dev_alloc(w, (void**)&w->pool, (pool_words ? pool_words : 512) * 4, "block pool")) !=
          VT_OK ||
      (st = dev_alloc(w, (void**)&w->atlas, atlas_words * 4, "model atlas")) != VT_OK) {
    free_world(w);
    return st;
  }
  w->pool_words = pool_words;
  w->pool_capacity = pool_words ? pool_words : 512;
  w->chunk_mirror = new (std::nothrow) uint32_t[nchunks];
  if (!w->chunk_mirror) {
    free_world(w);
    return fail(VT_ERR_OUT_OF_MEMORY, "vt_world_create_brickmap: host chunk table");
  }
  std::memcpy(w->chunk_mirror, chunks, nchunks * 4);
  vt::scan_brickmap_box(chunks, pool, dim, w);
  hipError_t e = hipMemcpy(w->chunks, chunks, nchunks * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && pool_words)
    e = hipMemcpy(w->pool, pool, pool_words * 4, hipMemcpyHostToDevice);
  else if (e == hipSuccess)  // the spare block edits allocate first starts zero
    e = hipMemset(w->pool, 0, 512 * 4);
  if (e == hipSuccess) e = hipMemcpy(w->atlas, atlas_rgba8, atlas_words * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_world(w);
    return hip_fail(e, "vt_world_create_brickmap: upload");
  }
  if ((dim & (dim - 1)) == 0) {  // power-of-two map: occupancy bytes for the fast loop
    const size_t n = (size_t)dim * dim * dim;
    if ((st = dev_alloc(w, (void**)&w->occ, n, "block occupancy")) != VT_OK ||
        (st = dev_alloc(w, (void**)&w->brick_bits, (size_t)vt::brick_bit_words(dim) * 8,
                        "block brick bits")) != VT_OK ||
        (st = alloc_freemap(w)) != VT_OK) {
      free_world(w);
      return st;
    }
    e = vt::launch_build_occupancy(vt::brickmap_view(w), w->occ, nullptr);
    if (e == hipSuccess) e = vt::rebuild_freemap(w, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = vt::finish_freemap(w);
    if (e != hipSuccess) {
      free_world(w);
      return hip_fail(e, "vt_world_create_brickmap: occupancy");
    }
  }
  *out = w;
  return VT_OK;
}

int vt_world_destroy(vt_world* world) {
  if (!world) return fail(VT_ERR_NULL_POINTER, "vt_world_destroy: world is NULL");
  (void)hipSetDevice(world->ctx->device);
  free_world(world);
  return VT_OK;
}

int vt_world_info(const vt_world* world, uint32_t* dim, size_t* device_bytes) {
  if (!world) return fail(VT_ERR_NULL_POINTER, "vt_world_info: world is NULL");
  if (dim) *dim = world->dim;
  if (device_bytes) *device_bytes = world->bytes;
  return VT_OK;
}

int vt_free_map_used(const vt_world* world, const vt_render_params* params,
                     const vt_frame* frame, int* used) {
  if (!world || !params || !frame || !used)
    return fail(VT_ERR_NULL_POINTER, "vt_free_map_used: world, params, frame, used required");
  *used = 0;
  int st = vt::check_params(*params, world, "vt_free_map_used");
  if (st == VT_OK) st = vt::check_frame(frame, params->flags, "vt_free_map_used");
  if (st != VT_OK) return st;
  *used = vt::render_free_map(world, *params, *frame) ? 1 : 0;
  return VT_OK;
}

int vt_render(vt_ctx* ctx, const vt_world* world, const vt_camera* camera,
              const vt_render_params* params, const vt_frame* frame, void* stream) {
  if (!ctx || !world || !camera || !params || !frame)
    return fail(VT_ERR_NULL_POINTER, "vt_render: ctx, world, camera, params, frame required");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_render: world from another ctx");
  const vt_render_params& p = *params;
  int st = vt::check_image_size(p.width, p.height, "vt_render");
  if (st == VT_OK) st = vt::check_params(p, world, "vt_render");
  if (st == VT_OK) st = vt::check_frame(frame, p.flags, "vt_render");
  if (st != VT_OK) return st;
  VT_HIP(hipSetDevice(ctx->device));
  const vt::ViewConsts v = vt::view_consts(camera, p.width, p.height);
  const vt::ShadeConsts sc = vt::shade_consts(p);
  const vt::FrameOut o = vt::frame_out(frame);
  // Tiles per wave: 16x4 pixels for frames that fill the GPU (at least 8
  // blocks of 4 waves per CU: 1080p 0.766 vs 0.769 ms for 8x8 and 0.780 for
  // 32x2; the engine's own frame 0.354 vs 0.365 ms), 8x8 below that (C3's
  // 512x512 views: 0.1165 vs 0.1185 ms per view).
  const uint32_t blocks_8x8 = ((p.width + 15u) / 16u) * ((p.height + 15u) / 16u);
  vt::TileShape tile{blocks_8x8 >= 8u * (uint32_t)ctx->compute_units ? 4u : 3u};
  hipError_t e;
  const bool free_map = vt::render_free_map(world, p, *frame);
  if (world->kind == WorldKind::Dense) {
    vt::DenseWorld dw = vt::dense_view(world);
    if (!free_map) dw.free_bytes = 0;
    e = vt::launch_render_dense(dw, p.traversal, v, sc, o, frame->schedule, tile,
                                (uint32_t)ctx->compute_units, (hipStream_t)stream);
  } else {
    vt::BrickmapWorld bw = vt::brickmap_view(world);
    if (!free_map) bw.free_bytes = 0;
    e = vt::launch_render_brickmap(bw, v, sc, o, frame->schedule, tile,
                                   (uint32_t)ctx->compute_units, (hipStream_t)stream);
  }
  if (e != hipSuccess) return hip_fail(e, "vt_render: launch");
  return VT_OK;
}

}  // extern "C"
