// vt_procgen.hip -- the reference world generator, written into a brickmap
// world on the device.
//
// procgen(dim, world, offsetX, offsetY) (src/procgen.zig:6-53, place_tree
// :55-70), which Game.init runs on the engine's map (game.zig:96-97), writes
// through VoxelBrickmap.set/get in a fixed order: a base layer of stone
// (y < 16 in every column), then per column in x-major order the terrain up
// to a noise height, a decoration on top and now and then a tree, drawing
// from one LCG whose draw count per column depends on what was drawn.  The
// work is split by what each part depends on:
//
//   1. k_procgen_heights (one lane per column): the column height
//      vh = trunc(max(fbm((ox + x) / 10, (oy + z) / 10) * dim * 0.1, 0)),
//      fbm = FastNoiseLite's fBm of 2D OpenSimplex2 noise with znoise's
//      generator defaults (seed 1337, frequency 0.01, 3 octaves, lacunarity
//      2, gain 0.5), in binary32 with every operation rounded once.
//   2. The host walk over the columns, in the generator's order, O(1) per
//      column: a column's terrain takes 2 vh + 1 draws (vh <= 16, base
//      layer materials only) or vh + 17 draws, skipped with the LCG's
//      affine jump (state after n draws = A_n s + C_n mod 2^32); after them
//      come at most a few decision draws (decoration 1 in 5, dirt 1 in 71,
//      tree 1 in 420) and the tree's own draws.  A column is cut short
//      (procgen.zig:37-38) when an earlier tree left a voxel at (x, vh, z).
//      The walk records each column's LCG state and decoration, the tree
//      voxels (last write wins) and allocates the chunks in the order the
//      generator's set() calls first touch them, so the chunk table and
//      pool come out word for word as the reference's.
//   3. k_procgen_columns (one wave per 8x8 chunk column): each lane
//      re-draws its column's terrain materials from the recorded state and
//      writes the base layer, the terrain and the decoration.
//   4. k_procgen_trees: the tree voxels at or above the terrain height of
//      their column.  A tree planted at column (x, z) writes only columns
//      (x..x+2, z..z+2), which the generator visits at or after (x, z); the
//      terrain of such a later column overwrote whatever tree voxels lay
//      below its height, and nothing else writes above it.
//   5. The block-occupancy bytes of the fast traversal are rebuilt.
#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "vt_internal.hpp"

namespace vt {
namespace {

constexpr uint32_t kSolid = 1u << 28;  // Voxel.is_solid (voxel.zig:7-11)
constexpr uint32_t kLcgMul = 1103515245u, kLcgAdd = 12345u, kLcgSeed = 0x46AE4Fu;
constexpr uint32_t kPrimeX = 501125321u, kPrimeY = 1136930381u;

// FastNoiseLite's 2D gradient pairs (gx, gy) = (sin a, cos a): pairs 0..119
// cycle five times through the 24 directions a = 7.5 + 15 m degrees, pairs
// 120..127 are the 8 directions a = 22.5 + 45 j degrees.  kSinDeg[k] =
// sin(7.5 + 15 k degrees), k = 0..5; the other sines follow by symmetry.
constexpr float kSinDeg[6] = {0.130526192220052f, 0.38268343236509f, 0.608761429008721f,
                              0.793353340291235f, 0.923879532511287f, 0.99144486137381f};

constexpr float sin_dir(int m) {  // sin(7.5 + 15 m degrees), m = 0..23
  return m < 6 ? kSinDeg[m] : m < 12 ? kSinDeg[11 - m] : m < 18 ? -kSinDeg[m - 12] : -kSinDeg[23 - m];
}

struct Grad2 {
  float v[256];
};

constexpr Grad2 make_grad2() {
  Grad2 t{};
  for (int p = 0; p < 128; p++) {
    const int m = p < 120 ? p % 24 : 1 + 3 * (p - 120);
    t.v[2 * p] = sin_dir(m);
    t.v[2 * p + 1] = sin_dir((m + 6) % 24);  // cos a = sin(a + 90 degrees)
  }
  return t;
}

__constant__ Grad2 kGrad2 = make_grad2();

__device__ __forceinline__ int fnl_floor(float f) { return f >= 0 ? (int)f : (int)f - 1; }

__device__ __forceinline__ float fnl_grad(int seed, uint32_t xp, uint32_t yp, float xd, float yd) {
  uint32_t h = ((uint32_t)seed ^ xp ^ yp) * 0x27d4eb2du;
  h ^= (uint32_t)((int32_t)h >> 15);
  h &= 254u;
  return xd * kGrad2.v[h] + yd * kGrad2.v[h | 1u];
}

// One octave: 2D simplex noise at skewed coordinates (x, y), the sum of the
// three corners of the containing triangle, each (0.5 - r^2)^4 (g . d).
__device__ float fnl_simplex2(int seed, float x, float y) {
  constexpr float SQRT3 = 1.7320508075688772935274463415059f;
  constexpr float G2 = (3 - SQRT3) / 6;
  constexpr float K1 = (float)(2 * (1 - 2 * G2) * (1 / G2 - 2));
  constexpr float K2 = (float)(-2 * (1 - 2 * G2) * (1 - 2 * G2));
  const int i = fnl_floor(x), j = fnl_floor(y);
  const float xi = x - (float)i, yi = y - (float)j;
  const float t = (xi + yi) * G2;
  const float x0 = xi - t, y0 = yi - t;
  const uint32_t ip = (uint32_t)i * kPrimeX, jp = (uint32_t)j * kPrimeY;
  const float a = 0.5f - x0 * x0 - y0 * y0;
  const float n0 = a <= 0 ? 0.0f : (a * a) * (a * a) * fnl_grad(seed, ip, jp, x0, y0);
  // the far corner (i+1, j+1): its 0.5 - r^2 is an affine function of t and a
  const float c = K1 * t + (K2 + a);
  float n2 = 0.0f;
  if (c > 0) {
    const float x2 = x0 + (2 * G2 - 1), y2 = y0 + (2 * G2 - 1);
    n2 = (c * c) * (c * c) * fnl_grad(seed, ip + kPrimeX, jp + kPrimeY, x2, y2);
  }
  // the middle corner: (i, j+1) above the diagonal, (i+1, j) below
  const bool upper = y0 > x0;
  const float x1 = upper ? x0 + G2 : x0 + (G2 - 1);
  const float y1 = upper ? y0 + (G2 - 1) : y0 + G2;
  const float b = 0.5f - x1 * x1 - y1 * y1;
  float n1 = 0.0f;
  if (b > 0)
    n1 = (b * b) * (b * b) *
         fnl_grad(seed, upper ? ip : ip + kPrimeX, upper ? jp + kPrimeY : jp, x1, y1);
  return (n0 + n1 + n2) * 99.83685446303647f;
}

// fnlGetNoise2D of FnlGenerator{ .fractal_type = .fbm } (procgen.zig:7).
__device__ float fnl_fbm2(float x, float y) {
  constexpr float SQRT3 = (float)1.7320508075688772935274463415059;
  constexpr float F2 = 0.5f * (SQRT3 - 1);
  constexpr float kBounding = 1.0f / (1.0f + 0.5f + 0.25f);  // 1 / sum of the octave gains
  x *= 0.01f;  // frequency
  y *= 0.01f;
  const float t = (x + y) * F2;  // OpenSimplex2 skew
  x += t;
  y += t;
  float sum = 0.0f, amp = kBounding;
  int seed = 1337;
  for (int o = 0; o < 3; o++) {
    const float noise = fnl_simplex2(seed++, x, y);
    sum += noise * amp;
    const float m = noise + 1 < 2 ? noise + 1 : 2;
    amp *= 1.0f + 0.0f * (m * 0.5f - 1.0f);  // weighted strength 0
    x *= 2.0f;                               // lacunarity
    y *= 2.0f;
    amp *= 0.5f;  // gain
  }
  return sum;
}

// Column heights, column index x * dim + z.
__global__ __launch_bounds__(256) void k_procgen_heights(uint32_t dim, float ox, float oy,
                                                         uint32_t* __restrict__ vh) {
  const uint32_t col = blockIdx.x * 256u + threadIdx.x;
  if (col >= dim * dim) return;
  const uint32_t x = col / dim, z = col % dim;
  const float val = fnl_fbm2((ox + (float)x) / 10.0f, (oy + (float)z) / 10.0f);
  const float h = val * (float)dim * 0.1f;
  vh[col] = (uint32_t)(h > 0.0f ? h : 0.0f);
}

__device__ __forceinline__ uint32_t lcg_next(uint32_t& s) {
  s = s * kLcgMul + kLcgAdd;
  return s;
}

// One wave per chunk column (cx, cz): lane = (x & 7) + 8 (z & 7).  Every
// chunk a lane writes was allocated by the host walk.
__global__ __launch_bounds__(64) void k_procgen_columns(uint32_t dim,
                                                        const uint32_t* __restrict__ chunks,
                                                        uint32_t* __restrict__ pool,
                                                        const uint32_t* __restrict__ vh,
                                                        const uint32_t* __restrict__ state,
                                                        const uint32_t* __restrict__ decor) {
  const uint32_t M = dim / 8u;
  const uint32_t cx = blockIdx.x / M, cz = blockIdx.x % M;
  const uint32_t lx = threadIdx.x & 7u, lz = threadIdx.x >> 3;
  const uint32_t x = cx * 8u + lx, z = cz * 8u + lz;
  const uint32_t col = x * dim + z;
  const uint32_t h = vh[col];
  uint32_t s = state[col];
  const uint32_t top = h > 16u ? h : 16u;
  const uint32_t lane_off = lx + 64u * lz;
  for (uint32_t y = 0; y < top; y++) {
    uint32_t w;
    if (y < h) {  // procgen.zig:25-33
      w = 21u + lcg_next(s) % 3u;
      if (y <= 15u)
        w = 25u + lcg_next(s) % 3u;
      else if (y == h - 1u)
        w = lcg_next(s) % 6u;
      w |= kSolid;
    } else {  // procgen.zig:10-19
      w = 13u | kSolid;
    }
    const uint32_t e = chunks[cx + M * ((y >> 3) + M * cz)];
    pool[(size_t)(e - 1u) * 512u + lane_off + 8u * (y & 7u)] = w;
  }
  const uint32_t d = decor[col];
  if (d != 0u) {
    const uint32_t e = chunks[cx + M * ((h >> 3) + M * cz)];
    pool[(size_t)(e - 1u) * 512u + lane_off + 8u * (h & 7u)] = d;
  }
}

struct TreeVoxel {
  uint64_t addr;  // pool word
  uint32_t value;
  uint32_t pad;
};

__global__ __launch_bounds__(256) void k_procgen_trees(uint32_t* __restrict__ pool,
                                                      const TreeVoxel* __restrict__ tv,
                                                      uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) pool[tv[i].addr] = tv[i].value;
}

// The host walk (steps 2 and 4 of the header).
struct Walk {
  uint32_t dim, M;
  std::vector<uint32_t> mirror;  // chunk table being allocated
  uint64_t used = 0;             // blocks allocated
  std::vector<uint32_t> cols;  // [col] LCG state at the column, [ncol + col] decoration
  std::unordered_map<uint64_t, uint32_t> trees;  // voxel x + dim (y + dim z) -> word
  std::vector<uint8_t> tree_col;                 // column x * dim + z holds a tree voxel
  uint32_t y_end = 16;                           // layers the generator wrote
  std::vector<uint32_t> jump_a, jump_c;          // n draws: s -> jump_a[n] s + jump_c[n]
  uint32_t bad_x = 0, bad_z = 0;

  uint64_t key(uint64_t x, uint64_t y, uint64_t z) const { return x + dim * (y + (uint64_t)dim * z); }
  void touch(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t& e = mirror[x / 8u + M * (y / 8u + M * (z / 8u))];
    if (e == 0) e = (uint32_t)(++used);
  }
  void tree_set(uint32_t x, uint32_t y, uint32_t z, uint32_t v) {
    touch(x, y, z);
    trees[key(x, y, z)] = v;
    tree_col[(size_t)x * dim + z] = 1;
    if (y + 1u > y_end) y_end = y + 1u;
  }
  static uint32_t next(uint32_t& s) {
    s = s * kLcgMul + kLcgAdd;
    return s;
  }

  // place_tree (procgen.zig:55-70); false if it would leave the map.
  bool place_tree(uint32_t& s, uint32_t x, uint32_t y, uint32_t z) {
    const uint32_t th = next(s) % 4u + 4u;
    if (x + 2u >= dim || z + 2u >= dim || y + th + 2u >= dim) {
      bad_x = x;
      bad_z = z;
      return false;
    }
    tree_set(x + 1u, y, z + 1u, 15u | kSolid);
    for (uint32_t o = 0; o < th; o++) tree_set(x + 1u, y + o, z + 1u, (14u + next(s) % 3u) | kSolid);
    for (uint32_t a = 0; a < 3; a++)
      for (uint32_t b = 0; b < 3; b++)
        for (uint32_t c = 0; c < 3; c++) tree_set(x + a, y + th + b, z + c, (18u + next(s) % 2u) | kSolid);
    return true;
  }

  bool run(const std::vector<uint32_t>& vh) {
    const uint32_t nmax = dim + 17u;  // most draws one column's terrain takes
    jump_a.assign(nmax + 1, 1u);
    jump_c.assign(nmax + 1, 0u);
    for (uint32_t n = 1; n <= nmax; n++) {
      jump_a[n] = jump_a[n - 1] * kLcgMul;
      jump_c[n] = jump_c[n - 1] * kLcgMul + kLcgAdd;
    }
    // base layer (procgen.zig:10-19): chunks (cx, 0, cz), (cx, 1, cz) in
    // first-touch order
    for (uint32_t cx = 0; cx < M; cx++)
      for (uint32_t cz = 0; cz < M; cz++) {
        touch(cx * 8u, 0u, cz * 8u);
        touch(cx * 8u, 8u, cz * 8u);
      }
    const size_t ncol = (size_t)dim * dim;
    cols.assign(2 * ncol, 0u);
    tree_col.assign((size_t)dim * dim, 0u);
    uint32_t s = kLcgSeed;
    for (uint32_t x = 0; x < dim; x++) {
      for (uint32_t z = 0; z < dim; z++) {
        const size_t col = (size_t)x * dim + z;
        const uint32_t h = vh[col];
        if (h >= dim) {  // the terrain would leave the map
          bad_x = x;
          bad_z = z;
          return false;
        }
        cols[col] = s;
        if (h <= 16u) {  // two draws per layer, then the closing draw
          s = jump_a[2u * h + 1u] * s + jump_c[2u * h + 1u];
          continue;
        }
        for (uint32_t y = 16u; y < h; y += 8u) touch(x, y, z);
        if (h + 1u > y_end) y_end = h + 1u;  // (with the decoration)
        s = jump_a[h + 17u] * s + jump_c[h + 17u];
        if (tree_col[col] && trees.count(key(x, h, z))) continue;  // procgen.zig:37-38
        uint32_t d = 0;
        if (next(s) % 5u == 0) d = 7u + next(s) % 5u;
        if (next(s) % 71u == 0) d = (7u + 5u) | kSolid;
        if (d) touch(x, h, z);
        cols[ncol + col] = d;
        if (next(s) % 420u == 0 && x < 500u && z < 500u && x > 5u && z > 5u)
          if (!place_tree(s, x, h, z)) return false;
        next(s);
      }
    }
    return true;
  }
};

}  // namespace
}  // namespace vt

extern "C" {

int vt_world_procgen(vt_ctx* ctx, vt_world* world, float offset_x, float offset_y,
                     void* stream) {
  using vt::fail;
  if (!ctx || !world) return fail(VT_ERR_NULL_POINTER, "vt_world_procgen: null argument");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_world_procgen");
  if (world->kind != WorldKind::Brickmap)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_world_procgen: needs a brickmap world");
  const uint32_t dim = world->dim;
  if (dim < 16u)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_procgen: dim %u is below the 16 layers of the base layer", dim);
  if (world->pool_words != 0)
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_procgen: the world already holds %zu blocks (procgen fills a new map)",
                world->pool_words / 512u);
  if (!(offset_x == offset_x) || !(offset_y == offset_y))
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_world_procgen: NaN offset");
  VT_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = (hipStream_t)stream;
  const size_t ncol = (size_t)dim * dim, M = dim / 8u, nchunks = M * M * M;

  struct Bufs {
    std::vector<void*> p;
    ~Bufs() {
      for (void* q : p) vt::dev_free(q);
    }
    hipError_t alloc(void** out, size_t bytes) {
      hipError_t e = hipMalloc(out, bytes ? bytes : 4);
      if (e == hipSuccess) p.push_back(*out);
      return e;
    }
  } bufs;
  uint32_t* d_vh;
  VT_HIP(bufs.alloc((void**)&d_vh, ncol * 4));
  hipLaunchKernelGGL(vt::k_procgen_heights, dim3((uint32_t)((ncol + 255) / 256)), dim3(256), 0, s,
                     dim, offset_x, offset_y, d_vh);
  VT_HIP(hipGetLastError());
  std::vector<uint32_t> vh(ncol);
  VT_HIP(hipMemcpyAsync(vh.data(), d_vh, ncol * 4, hipMemcpyDeviceToHost, s));
  VT_HIP(hipStreamSynchronize(s));

  vt::Walk walk;
  walk.dim = dim;
  walk.M = (uint32_t)M;
  walk.mirror.assign(nchunks, 0u);
  if (!walk.run(vh))
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_world_procgen: column (%u, %u) writes outside the %u^3 map (its terrain, "
                "or its tree: place_tree writes up to x+2, z+2 and 9 layers up)",
                walk.bad_x, walk.bad_z, dim);
  std::vector<vt::TreeVoxel> tv;
  tv.reserve(walk.trees.size());
  for (const auto& kv : walk.trees) {
    const uint64_t x = kv.first % dim, y = kv.first / dim % dim, z = kv.first / dim / dim;
    if (y < vh[x * dim + z]) continue;  // under a later column's terrain
    const uint32_t e = walk.mirror[x / 8u + M * (y / 8u + M * (z / 8u))];
    tv.push_back({(uint64_t)(e - 1u) * 512u + x % 8u + 8u * (y % 8u + 8u * (z % 8u)), kv.second,
                  0u});
  }

  int st = vt::grow_pool(world, walk.used, s, "vt_world_procgen: pool growth");
  if (st != VT_OK) return st;  // nothing written: the map is still empty
  // one staging allocation: column states and decorations, then the tree voxels
  uint32_t* d_cols;
  VT_HIP(bufs.alloc((void**)&d_cols, 2 * ncol * 4 + tv.size() * sizeof(vt::TreeVoxel)));
  vt::TreeVoxel* d_tv = (vt::TreeVoxel*)(d_cols + 2 * ncol);
  // From here on the device map changes; the host chunk table and pool count
  // follow only once all of it succeeded.
  hipError_t e = hipMemcpyAsync(world->chunks, walk.mirror.data(), nchunks * 4,
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_cols, walk.cols.data(), 2 * ncol * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !tv.empty())
    e = hipMemcpyAsync(d_tv, tv.data(), tv.size() * sizeof(vt::TreeVoxel), hipMemcpyHostToDevice,
                       s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(vt::k_procgen_columns, dim3((uint32_t)(M * M)), dim3(64), 0, s, dim,
                       world->chunks, world->pool, d_vh, d_cols, d_cols + ncol);
    e = hipGetLastError();
  }
  if (e == hipSuccess && !tv.empty()) {
    hipLaunchKernelGGL(vt::k_procgen_trees, dim3((uint32_t)((tv.size() + 255) / 256)), dim3(256),
                       0, s, world->pool, d_tv, (uint32_t)tv.size());
    e = hipGetLastError();
  }
  // the empty map's occupancy bytes were zero; only the written layers change
  if (e == hipSuccess && world->occ)
    e = vt::launch_build_occupancy(vt::brickmap_view(world), world->occ, s,
                                   std::min(dim, (walk.y_end + 7u) & ~7u));
  if (e == hipSuccess) e = vt::rebuild_freemap(world, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    // back to the empty map the call was given: chunk table, written pool
    // blocks and occupancy bytes zero
    (void)hipMemsetAsync(world->chunks, 0, nchunks * 4, s);
    (void)hipMemsetAsync(world->pool, 0, walk.used * 512u * 4u, s);
    if (world->occ) (void)hipMemsetAsync(world->occ, 0, (size_t)dim * dim * dim, s);
    (void)hipStreamSynchronize(s);
    (void)hipGetLastError();
    return vt::hip_fail(e, "vt_world_procgen: device writes");
  }
  std::memcpy(world->chunk_mirror, walk.mirror.data(), nchunks * 4);
  world->pool_words = walk.used * 512u;
  vt::scan_brickmap_box(world->chunk_mirror, nullptr, dim, world);  // chunk granularity
  VT_HIP(vt::finish_freemap(world));
  return VT_OK;
}

}  // extern "C"
