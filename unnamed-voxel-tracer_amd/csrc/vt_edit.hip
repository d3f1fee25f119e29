// vt_edit.hip -- world edits, voxel reads and the centre-ray pick.
//
// Edits replace VoxelBrickmap.set (src/engine/voxel.zig:46-61, the CPU
// writing through a persistent mapping) and map_setVoxel (map.glsl:49-55);
// reads replace VoxelBrickmap.get / map_getVoxel (voxel.zig:63-70,
// map.glsl:38-47); the pick is terrain_edit.comp.glsl:11-17.
//
// A batch of edits means what calling set() once per edit, in order, means:
// chunks are allocated on first touch in edit order (pool block index =
// allocation count, voxel.zig:47-56), and of several edits of one voxel the
// last wins.  The host resolves that order (chunk allocation, last-writer
// dedup) and the device applies the surviving writes in one scatter:
//   brickmap: pool words and new chunk-table entries; the pool grows by
//             doubling like GpuBlockAllocator.alloc (gpu_block_allocator.zig:
//             20-30), new blocks zero;
//   dense:    Morton-ordered bytes, then the occupancy words of every touched
//             4^3 brick (one wave per brick: ballot over its 64 bytes) and
//             of every touched 16^3 region (ballot over its 64 brick words),
//             then the compact albedo index, when built, is recomputed in
//             place (its numbering changes with the occupancy).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

__global__ __launch_bounds__(256) void k_scatter_u32(uint32_t* __restrict__ dst,
                                                     const uint64_t* __restrict__ addr,
                                                     const uint32_t* __restrict__ val, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[addr[i]] = val[i];
}

// dst[addr[i]] = 0 for i < n (undoing k_scatter_u32 of new chunk entries).
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ dst,
                                                  const uint64_t* __restrict__ addr, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[addr[i]] = 0u;
}

__global__ __launch_bounds__(256) void k_scatter_u8(uint8_t* __restrict__ dst,
                                                    const uint64_t* __restrict__ addr,
                                                    const uint32_t* __restrict__ val, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[addr[i]] = (uint8_t)val[i];
}

__global__ __launch_bounds__(256) void k_scatter_nonzero_u8(uint8_t* __restrict__ dst,
                                                            const uint64_t* __restrict__ addr,
                                                            const uint32_t* __restrict__ val,
                                                            uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[addr[i]] = val[i] != 0u ? 1 : 0;
}

// One wave per listed word: bit l of out[word] = (in[word * 64 + l] != 0).
template <class T>
__global__ __launch_bounds__(256) void k_rebuild_mask(const T* __restrict__ in,
                                                      uint64_t* __restrict__ out,
                                                      const uint32_t* __restrict__ words,
                                                      uint32_t n) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  if (wave >= n) return;  // whole waves leave together
  const uint32_t word = words[wave];
  const uint64_t m = __ballot(in[(size_t)word * 64u + lane] != 0);
  if (lane == 0) out[word] = m;
}

__global__ void k_gather_dense(const DenseWorld w, const uint32_t* __restrict__ xyz, uint32_t n,
                               uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = w.vox[morton3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2])];
}

__global__ void k_gather_brickmap(const BrickmapWorld w, const uint32_t* __restrict__ xyz,
                                  uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = bm_block(w, (int)xyz[3 * i], (int)xyz[3 * i + 1], (int)xyz[3 * i + 2]);
}

// terrain_edit.comp.glsl:11-17: rayUV = 0, then the primary ray's
// normalize(C_view * vec4(rayUV, 1, 1)).xyz (4 components), AABB entry and
// EPSILON offset.
__device__ __forceinline__ void centre_ray(const ViewConsts& v, uint32_t dim, float e[3],
                                           float d[3]) {
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float s = v.view[i] * 0.0f;
    s = s + v.view[4 + i] * 0.0f;
    s = s + v.view[8 + i] * 1.0f;
    s = s + v.view[12 + i] * 1.0f;
    r[i] = s;
  }
  float q = r[0] * r[0];
  q = q + r[1] * r[1];
  q = q + r[2] * r[2];
  q = q + r[3] * r[3];
  const float len = sqrtf(q);
  d[0] = r[0] / len;
  d[1] = r[1] / len;
  d[2] = r[2] / len;
  float tn, tf;
  const float D = (float)dim;
  intersect_aabb(v.pos[0], v.pos[1], v.pos[2], d[0], d[1], d[2], 0.0f, 0.0f, 0.0f, D, D, D, tn,
                 tf);
  const float tc = fmaxf(tn, 0.0f);
#pragma unroll
  for (int i = 0; i < 3; i++) e[i] = (v.pos[i] + d[i] * tc) - 0.001f;
}

// HitInfo of the pick (map.glsl:127 on a hit, :167 on a miss) and the cells
// of the shader's disabled edit body (terrain_edit.comp.glsl:21,23).
__device__ __forceinline__ vt_pick_result pick_result(bool hit, uint32_t data, int face,
                                                      float hx, float hy, float hz, int steps) {
  vt_pick_result res = {};
  res.steps = (uint32_t)steps;
  if (hit) {
    res.data = data;
    res.face = (uint32_t)face;
    res.hit_pos[0] = hx;
    res.hit_pos[1] = hy;
    res.hit_pos[2] = hz;
    face_normal(face, res.normal[0], res.normal[1], res.normal[2]);
  } else {
    res.hit_pos[0] = res.hit_pos[1] = res.hit_pos[2] = -1.0f;
  }
  for (int i = 0; i < 3; i++) {
    res.place[i] = f2i_sat(res.hit_pos[i] + res.normal[i]) >> 3;
    res.remove[i] = f2i_sat(res.hit_pos[i]) >> 3;
  }
  return res;
}

__global__ void k_pick_dense(const DenseWorld w, ViewConsts v, int max_steps,
                             vt_pick_result* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float e[3], d[3];
  centre_ray(v, w.dim, e, d);
  DenseRay ray;
  FlatLookup lk;
  DenseHit h;
  uint32_t l1f = 0, l0f = 0, leaf = 0;
  if (!dense_begin<false>(ray, lk, w, e[0], e[1], e[2], d[0], d[1], d[2], max_steps, h, l1f, l0f))
    dense_run<false>(ray, lk, w, h, l1f, l0f);
  const uint32_t data = h.hit ? w.palette[lk.material(w, h.code, h.v, leaf)] : 0u;
  *out = pick_result(h.hit, data, h.face, h.hx, h.hy, h.hz, h.steps);
}

__global__ void k_pick_brickmap(const BrickmapWorld w, ViewConsts v, int max_steps,
                                vt_pick_result* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float e[3], d[3];
  centre_ray(v, w.dim, e, d);
  GeneralHit g;
  uint32_t chunk_reads = 0, pool_reads = 0;
  trace_general<false>(w, e[0], e[1], e[2], d[0], d[1], d[2], max_steps, g, chunk_reads,
                       pool_reads);
  *out = pick_result(g.data != 0u, g.data, g.face, g.hx, g.hy, g.hz, g.steps);
}

// Grows the block pool to at least `blocks` blocks (doubling), keeping the
// contents and zeroing the new blocks.
int grow_pool(vt_world* w, uint64_t blocks, hipStream_t s, const char* who) {
  const uint64_t cap = w->pool_capacity / 512u;
  if (blocks <= cap) return VT_OK;
  uint64_t ncap = cap ? cap : 1;
  while (ncap < blocks) ncap *= 2;
  // An optional budget for the block pool (bytes): a growth past it fails
  // like a failed allocation, before anything changes.
  if (const char* lim = std::getenv("VT_POOL_LIMIT_BYTES")) {
    const unsigned long long limit = std::strtoull(lim, nullptr, 10);
    if (limit && ncap * 512u * 4u > limit)
      return fail(VT_ERR_OUT_OF_MEMORY, "%s: the pool would grow to %llu bytes, past "
                  "VT_POOL_LIMIT_BYTES=%llu", who, (unsigned long long)(ncap * 512u * 4u), limit);
  }
  uint32_t* np = nullptr;
  hipError_t e = hipMalloc((void**)&np, ncap * 512u * 4u);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return hip_fail(e, who);
  }
  e = hipMemcpyAsync(np, w->pool, cap * 512u * 4u, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(np + cap * 512u, 0, (ncap - cap) * 512u * 4u, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    dev_free(np);
    return hip_fail(e, who);
  }
  dev_free(w->pool);
  w->pool = np;
  w->bytes += (ncap - cap) * 512u * 4u;
  w->pool_capacity = ncap * 512u;
  return VT_OK;
}

}  // namespace vt

namespace {

using vt::fail;
using vt::hip_fail;

// Device staging for one call; freed after the call's stream work is done.
struct Staging {
  std::vector<void*> bufs;
  ~Staging() {
    for (void* p : bufs) vt::dev_free(p);
  }
  template <class T>
  hipError_t upload(const std::vector<T>& v, T** out, hipStream_t s) {
    *out = nullptr;
    if (v.empty()) return hipSuccess;
    hipError_t e = hipMalloc((void**)out, v.size() * sizeof(T));
    if (e != hipSuccess) return e;
    bufs.push_back(*out);
    return hipMemcpyAsync(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

uint32_t part1by2_host(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0xFF0000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  return (v | (v << 2)) & 0x09249249u;
}

uint32_t morton_code(uint32_t x, uint32_t y, uint32_t z) {
  return part1by2_host(x) | (part1by2_host(y) << 1) | (part1by2_host(z) << 2);
}

uint32_t blocks_for(uint64_t n, uint32_t per_block) {
  return (uint32_t)((n + per_block - 1) / per_block);
}

int check_coords(const vt_world* w, const uint32_t* xyz, uint32_t n, const char* who) {
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++)
      if (xyz[3 * (size_t)i + k] >= w->dim)
        return fail(VT_ERR_INVALID_ARGUMENT, "%s: voxel %u coordinate %u outside [0, %u)", who, i,
                    xyz[3 * (size_t)i + k], w->dim);
  return VT_OK;
}

// The box of the nonempty voxels grows with every voxel set nonzero (a
// cleared voxel leaves it as it was: still a box around all of them).
static void grow_box_by(vt_world* w, const uint32_t* xyz, const uint32_t* values, uint32_t n) {
  for (uint32_t i = 0; i < n; i++)
    if (values[i] != 0u)
      vt::grow_box(w, (int32_t)xyz[3 * (size_t)i], (int32_t)xyz[3 * (size_t)i + 1],
                   (int32_t)xyz[3 * (size_t)i + 2]);
}

int set_brickmap(vt_world* w, const uint32_t* xyz, const uint32_t* values, uint32_t n,
                 hipStream_t s) {
  const uint32_t M = w->dim / 8u;
  std::vector<std::pair<uint64_t, uint32_t>> new_chunks;
  std::unordered_map<uint64_t, uint32_t> last;  // pool word -> position in `writes`
  std::vector<uint64_t> addr, oaddr;            // pool word, occupancy byte (Morton)
  std::vector<uint32_t> val;
  // chunks this call allocates: kept here and entered into the host chunk
  // table only once the device holds them, so a failed call leaves the
  // world as it was
  std::unordered_map<uint64_t, uint32_t> fresh;
  uint64_t used = w->pool_words / 512u;
  last.reserve(n);
  grow_box_by(w, xyz, values, n);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t x = xyz[3 * (size_t)i], y = xyz[3 * (size_t)i + 1], z = xyz[3 * (size_t)i + 2];
    const uint64_t ci = (uint64_t)(x / 8u) + M * ((uint64_t)(y / 8u) + (uint64_t)M * (z / 8u));
    uint32_t entry = w->chunk_mirror[ci];
    if (entry == 0) {
      auto f = fresh.find(ci);
      if (f != fresh.end()) {
        entry = f->second;
      } else {  // get_block_for_chunk: allocate on first touch
        if (used >= 0xFFFFFFFEull) return fail(VT_ERR_OUT_OF_MEMORY, "vt_world_set_voxels: pool");
        entry = (uint32_t)(++used);
        fresh.emplace(ci, entry);
        new_chunks.emplace_back(ci, entry);
      }
    }
    const uint64_t a = (uint64_t)(entry - 1u) * 512u + (x % 8u) + 8u * ((y % 8u) + 8u * (z % 8u));
    auto it = last.find(a);
    if (it == last.end()) {
      last.emplace(a, (uint32_t)addr.size());
      addr.push_back(a);
      val.push_back(values[i]);
      if (w->occ) oaddr.push_back(morton_code(x, y, z));
    } else {
      val[it->second] = values[i];
    }
  }
  int st = vt::grow_pool(w, used, s, "vt_world_set_voxels: pool growth");
  if (st != VT_OK) return st;  // nothing written: the world is unchanged
  std::vector<uint64_t> caddr;
  std::vector<uint32_t> cval;
  for (auto& c : new_chunks) {
    caddr.push_back(c.first);
    cval.push_back(c.second);
  }
  Staging stg;
  uint64_t *d_addr, *d_caddr, *d_oaddr;
  uint32_t *d_val, *d_cval;
  VT_HIP(stg.upload(addr, &d_addr, s));
  VT_HIP(stg.upload(val, &d_val, s));
  VT_HIP(stg.upload(caddr, &d_caddr, s));
  VT_HIP(stg.upload(cval, &d_cval, s));
  VT_HIP(stg.upload(oaddr, &d_oaddr, s));
  if (!addr.empty())
    hipLaunchKernelGGL(vt::k_scatter_u32, dim3(blocks_for(addr.size(), 256)), dim3(256), 0, s,
                       w->pool, d_addr, d_val, (uint32_t)addr.size());
  if (!oaddr.empty())
    hipLaunchKernelGGL(vt::k_scatter_nonzero_u8, dim3(blocks_for(oaddr.size(), 256)), dim3(256),
                       0, s, w->occ, d_oaddr, d_val, (uint32_t)oaddr.size());
  if (!caddr.empty())
    hipLaunchKernelGGL(vt::k_scatter_u32, dim3(blocks_for(caddr.size(), 256)), dim3(256), 0, s,
                       w->chunks, d_caddr, d_cval, (uint32_t)caddr.size());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && !oaddr.empty()) e = vt::rebuild_freemap(w, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    // Undo what may have reached the device: new chunk entries back to 0
    // and the new blocks zero again (edits allocate zero blocks), so the
    // device agrees with the unchanged host chunk table and pool count.
    if (!caddr.empty())
      hipLaunchKernelGGL(vt::k_fill_u32, dim3(blocks_for(caddr.size(), 256)), dim3(256), 0, s,
                         w->chunks, d_caddr, (uint32_t)caddr.size());
    const uint64_t b0 = w->pool_words / 512u;
    if (used > b0)
      (void)hipMemsetAsync(w->pool + b0 * 512u, 0, (used - b0) * 512u * 4u, s);
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    return hip_fail(e, "vt_world_set_voxels: scatter");
  }
  for (auto& c : new_chunks) w->chunk_mirror[c.first] = c.second;
  w->pool_words = used * 512u;
  if (!oaddr.empty()) VT_HIP(vt::finish_freemap(w));
  return VT_OK;
}

int set_dense(vt_world* w, const uint32_t* xyz, const uint32_t* values, uint32_t n,
              hipStream_t s) {
  for (uint32_t i = 0; i < n; i++)
    if (values[i] > 255u)
      return fail(VT_ERR_INVALID_ARGUMENT, "vt_world_set_voxels: dense value %u > 255 (edit %u)",
                  values[i], i);
  std::unordered_map<uint64_t, uint32_t> last;
  std::vector<uint64_t> addr;
  std::vector<uint32_t> val, bricks, regions, regions8;
  last.reserve(n);
  grow_box_by(w, xyz, values, n);
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t code =
        morton_code(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]);
    auto it = last.find(code);
    if (it == last.end()) {
      last.emplace(code, (uint32_t)addr.size());
      addr.push_back(code);
      val.push_back(values[i]);
      bricks.push_back(code >> 6);
      regions.push_back(code >> 12);
      regions8.push_back(code >> 15);
    } else {
      val[it->second] = values[i];
    }
  }
  std::sort(bricks.begin(), bricks.end());
  bricks.erase(std::unique(bricks.begin(), bricks.end()), bricks.end());
  std::sort(regions.begin(), regions.end());
  regions.erase(std::unique(regions.begin(), regions.end()), regions.end());
  std::sort(regions8.begin(), regions8.end());
  regions8.erase(std::unique(regions8.begin(), regions8.end()), regions8.end());
  Staging stg;
  uint64_t* d_addr;
  uint32_t *d_val, *d_bricks, *d_regions, *d_regions8;
  VT_HIP(stg.upload(addr, &d_addr, s));
  VT_HIP(stg.upload(val, &d_val, s));
  VT_HIP(stg.upload(bricks, &d_bricks, s));
  VT_HIP(stg.upload(regions, &d_regions, s));
  VT_HIP(stg.upload(regions8, &d_regions8, s));
  if (!addr.empty()) {
    hipLaunchKernelGGL(vt::k_scatter_u8, dim3(blocks_for(addr.size(), 256)), dim3(256), 0, s,
                       w->vox, d_addr, d_val, (uint32_t)addr.size());
    hipLaunchKernelGGL((vt::k_rebuild_mask<uint8_t>), dim3(blocks_for(bricks.size(), 4)),
                       dim3(256), 0, s, w->vox, w->l0, d_bricks, (uint32_t)bricks.size());
    hipLaunchKernelGGL((vt::k_rebuild_mask<uint64_t>), dim3(blocks_for(regions.size(), 4)),
                       dim3(256), 0, s, w->l0, w->l1, d_regions, (uint32_t)regions.size());
    VT_HIP(vt::launch_build_l1b(w->l0, (uint64_t)w->dim * w->dim * w->dim / 64u, d_regions8,
                                (uint32_t)regions8.size(), w->l1b, s));
    VT_HIP(vt::rebuild_freemap(w, s));
  }
  VT_HIP(hipGetLastError());
  if (w->prefix && !addr.empty()) {  // keep the compact index current
    const uint64_t nwords = (uint64_t)w->dim * w->dim * w->dim / 64u;
    const uint64_t ntiles = (nwords + 1023) / 1024;
    uint32_t* scratch = nullptr;
    VT_HIP(hipMalloc((void**)&scratch, (ntiles + 1) * 4));
    stg.bufs.push_back(scratch);
    VT_HIP(vt::launch_compact_prefix(w->l0, nwords, w->prefix, scratch, s));
    uint32_t total = 0;
    VT_HIP(hipMemcpyAsync(&total, scratch + ntiles, 4, hipMemcpyDeviceToHost, s));
    VT_HIP(hipStreamSynchronize(s));
    w->n_occupied = total;
  }
  VT_HIP(hipStreamSynchronize(s));
  if (!addr.empty()) VT_HIP(vt::finish_freemap(w));
  return VT_OK;
}

}  // namespace

extern "C" {

int vt_world_set_voxels(vt_ctx* ctx, vt_world* world, const uint32_t* xyz,
                        const uint32_t* values, uint32_t n, void* stream) {
  if (!ctx || !world || (n && (!xyz || !values)))
    return fail(VT_ERR_NULL_POINTER, "vt_world_set_voxels: null argument");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_world_set_voxels");
  int st = check_coords(world, xyz, n, "vt_world_set_voxels");
  if (st != VT_OK || n == 0) return st;
  VT_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = (hipStream_t)stream;
  return world->kind == WorldKind::Dense ? set_dense(world, xyz, values, n, s)
                                         : set_brickmap(world, xyz, values, n, s);
}

int vt_world_get_voxels(vt_ctx* ctx, const vt_world* world, const uint32_t* xyz, uint32_t n,
                        uint32_t* values, void* stream) {
  if (!ctx || !world || (n && (!xyz || !values)))
    return fail(VT_ERR_NULL_POINTER, "vt_world_get_voxels: null argument");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_world_get_voxels");
  int st = check_coords(world, xyz, n, "vt_world_get_voxels");
  if (st != VT_OK || n == 0) return st;
  VT_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = (hipStream_t)stream;
  // the context's staging: coordinates, then values, kCtxStageVoxels at a time
  uint32_t* d_xyz = (uint32_t*)ctx->stage;
  uint32_t* d_out = d_xyz + 3u * (size_t)kCtxStageVoxels;
  for (uint32_t i0 = 0; i0 < n; i0 += kCtxStageVoxels) {
    const uint32_t m = std::min(n - i0, kCtxStageVoxels);
    VT_HIP(hipMemcpyAsync(d_xyz, xyz + 3 * (size_t)i0, 12 * (size_t)m, hipMemcpyHostToDevice, s));
    if (world->kind == WorldKind::Dense)
      hipLaunchKernelGGL(vt::k_gather_dense, dim3(blocks_for(m, 256)), dim3(256), 0, s,
                         vt::dense_view(world), d_xyz, m, d_out);
    else
      hipLaunchKernelGGL(vt::k_gather_brickmap, dim3(blocks_for(m, 256)), dim3(256), 0, s,
                         vt::brickmap_view(world), d_xyz, m, d_out);
    VT_HIP(hipGetLastError());
    VT_HIP(hipMemcpyAsync(values + i0, d_out, 4 * (size_t)m, hipMemcpyDeviceToHost, s));
    VT_HIP(hipStreamSynchronize(s));
  }
  return VT_OK;
}

int vt_world_read_brickmap(vt_ctx* ctx, const vt_world* world, uint32_t* chunks,
                           uint32_t* pool, size_t pool_capacity_words, size_t* pool_words,
                           void* stream) {
  if (!ctx || !world || !pool_words)
    return fail(VT_ERR_NULL_POINTER, "vt_world_read_brickmap: null argument");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_world_read_brickmap");
  if (world->kind != WorldKind::Brickmap)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_world_read_brickmap: needs a brickmap world");
  *pool_words = world->pool_words;
  if (pool && pool_capacity_words < world->pool_words)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_read_brickmap: pool holds %zu words, the output room is %zu",
                world->pool_words, pool_capacity_words);
  VT_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = (hipStream_t)stream;
  const size_t M = world->dim / 8u;
  if (chunks)
    VT_HIP(hipMemcpyAsync(chunks, world->chunks, M * M * M * 4, hipMemcpyDeviceToHost, s));
  if (pool && world->pool_words)
    VT_HIP(hipMemcpyAsync(pool, world->pool, world->pool_words * 4, hipMemcpyDeviceToHost, s));
  VT_HIP(hipStreamSynchronize(s));
  return VT_OK;
}

int vt_world_read_free_map(vt_ctx* ctx, const vt_world* world, uint32_t* words,
                           size_t capacity_words, uint32_t* words_per_octant, int* valid,
                           void* stream) {
  if (!ctx || !world || !words_per_octant || !valid)
    return fail(VT_ERR_NULL_POINTER, "vt_world_read_free_map: null argument");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_world_read_free_map");
  const uint32_t wpo = vt::free_words_per_octant(world->dim);
  *words_per_octant = wpo;
  *valid = world->freemap && world->free_valid ? 1 : 0;
  if (!words || !world->freemap) return VT_OK;
  if (capacity_words < 8u * (size_t)wpo)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_read_free_map: the map holds %u words, the output room is %zu",
                8u * wpo, capacity_words);
  VT_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = (hipStream_t)stream;
  VT_HIP(hipMemcpyAsync(words, world->freemap, 8u * (size_t)wpo * 4u, hipMemcpyDeviceToHost, s));
  VT_HIP(hipStreamSynchronize(s));
  return VT_OK;
}

int vt_pick(vt_ctx* ctx, const vt_world* world, const vt_camera* camera, int32_t max_steps,
            vt_pick_result* result, void* stream) {
  if (!ctx || !world || !camera || !result)
    return fail(VT_ERR_NULL_POINTER, "vt_pick: null argument");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_pick");
  if (max_steps < 0 || max_steps > 65535)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_pick: max_steps %d outside [0, 65535]", max_steps);
  VT_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = (hipStream_t)stream;
  const vt::ViewConsts v = vt::view_consts(camera, 1, 1);
  vt_pick_result* d = (vt_pick_result*)ctx->stage;  // the context's staging
  if (world->kind == WorldKind::Dense) {
    vt::DenseWorld dw = vt::dense_view(world);
    dw.free_bytes = 0;  // the result reports traceMap's step count
    hipLaunchKernelGGL(vt::k_pick_dense, dim3(1), dim3(64), 0, s, dw, v, (int)max_steps, d);
  } else
    hipLaunchKernelGGL(vt::k_pick_brickmap, dim3(1), dim3(64), 0, s, vt::brickmap_view(world), v,
                       (int)max_steps, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(result, d, sizeof *d, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "vt_pick");
  return VT_OK;
}

}  // extern "C"
