// vt_trace.hpp -- device-side REF voxel traversal for gfx950.
//
// Restates traceMap (assets/shaders/map.glsl:83-168) and its callers
// (primary.comp.glsl:23-69, secondary.comp.glsl:18-50, blit.fragment.glsl:30-31)
// in binary32 with the same operation order as the CPU restatement: built with
// -ffp-contract=off and correctly rounded division/sqrt, every multiply and add
// rounds once, so hits are bit-exact.
//
// Two traversal loops:
//  * DenseTracer: worlds where one grid voxel is one reference block whose
//    atlas model is a solid cube.  After the first lookup traceMap never leaves
//    the block level (stepSize 3), so the state is kept in block units
//    G = gridsCoords >> 3 with withinGridCoords still in x8 units; this is the
//    same recurrence, with pos >> 3 == G + (int(within) >> 3).
//  * The literal general loop (trace_general) with both levels, used for the
//    reference brickmap + atlas layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vt {

// ---------------------------------------------------------------------------
// binary32 helpers (same definitions as the CPU restatement)
// ---------------------------------------------------------------------------

// ivec3(float) at ray set-up: truncation toward zero, NaN -> 0, saturating.
__device__ __forceinline__ int f2i_sat(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (-2147483647 - 1);
  return (int)x;
}

__device__ __forceinline__ float len3(float x, float y, float z) {
  float s = x * x;
  s = s + y * y;
  s = s + z * z;
  return sqrtf(s);
}

// rgba8 UNORM store followed by a load (gbuffer.zig:10-13 formats).
__device__ __forceinline__ float unorm8(float x) {
  float c = fminf(fmaxf(x, 0.0f), 1.0f);
  float b = floorf(c * 255.0f + 0.5f);
  return b / 255.0f;
}

// normals[] (map.glsl:72-79), face id 1..6
__device__ __forceinline__ void face_normal(int face, float& nx, float& ny, float& nz) {
  nx = face == 1 ? -1.0f : (face == 2 ? 1.0f : 0.0f);
  ny = face == 3 ? -1.0f : (face == 4 ? 1.0f : 0.0f);
  nz = face == 5 ? -1.0f : (face == 6 ? 1.0f : 0.0f);
}

// ---------------------------------------------------------------------------
// per-view constants
// ---------------------------------------------------------------------------
struct ViewConsts {
  float pos[3];     // C_position.xyz
  float view[16];   // C_view, column-major
  float tan_half;   // tan(fov/2), evaluated once per view on the host
  uint32_t W, H;
};

// primary.comp.glsl:31-39
__device__ __forceinline__ void raygen(const ViewConsts& v, uint32_t px, uint32_t py, float& dx,
                                       float& dy, float& dz) {
  float u = (float)px / (float)v.W;
  float w = (float)py / (float)v.H;
  u = u * 2.0f - 1.0f;
  w = w * 2.0f - 1.0f;
  w = w * ((float)v.H / (float)v.W);
  u = u * v.tan_half;
  w = w * v.tan_half;
  float r[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float acc = v.view[0 * 4 + i] * u;
    acc = acc + v.view[1 * 4 + i] * w;
    acc = acc + v.view[2 * 4 + i] * 1.0f;
    acc = acc + v.view[3 * 4 + i] * 1.0f;
    r[i] = acc;
  }
  float s = r[0] * r[0];
  s = s + r[1] * r[1];
  s = s + r[2] * r[2];
  s = s + r[3] * r[3];
  float l = sqrtf(s);
  dx = r[0] / l;
  dy = r[1] / l;
  dz = r[2] / l;
}

// intersectAABB (map.glsl:21-29)
__device__ __forceinline__ void intersect_aabb(float ox, float oy, float oz, float dx, float dy,
                                               float dz, float x0, float y0, float z0, float x1,
                                               float y1, float z1, float& tn, float& tf) {
  float ax = (x0 - ox) / dx, bx = (x1 - ox) / dx;
  float ay = (y0 - oy) / dy, by = (y1 - oy) / dy;
  float az = (z0 - oz) / dz, bz = (z1 - oz) / dz;
  float t1x = fminf(ax, bx), t2x = fmaxf(ax, bx);
  float t1y = fminf(ay, by), t2y = fmaxf(ay, by);
  float t1z = fminf(az, bz), t2z = fmaxf(az, bz);
  tn = fmaxf(fmaxf(t1x, t1y), t1z);
  tf = fminf(fminf(t2x, t2y), t2z);
}

// traceEntities occlusion (map.glsl:172-201); the code after the return on
// map.glsl:201 is unreachable in the reference.
__device__ __forceinline__ bool trace_entities(float ox, float oy, float oz, float dx, float dy,
                                               float dz, float max_distance) {
  const float PX[5] = {256.f, 251.f, 253.f, 251.f, 257.f};
  const float PY[5] = {21.f, 21.f, 21.f, 21.f, 21.f};
  const float PZ[5] = {256.f, 259.f, 256.f, 256.f, 261.f};
  float prevD = __builtin_inff();
  int id = -1;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    float dist = len3(ox - PX[i], oy - PY[i], oz - PZ[i]);
    if (dist >= max_distance) continue;
    float tn, tf;
    intersect_aabb(ox, oy, oz, dx, dy, dz, PX[i], PY[i], PZ[i], PX[i] + 1.0f, PY[i] + 1.0f,
                   PZ[i] + 1.0f, tn, tf);
    if (tf >= tn && prevD >= tf) {
      id = i;
      prevD = tf;
    }
  }
  return id >= 0;  // the re-test on map.glsl:199-200 repeats a test that passed
}

// SkyDome2 (camera.glsl:11-19)
__device__ __forceinline__ void sky(const float sun[3], float rx, float ry, float rz, float& r,
                                    float& g, float& b) {
  float ls = len3(sun[0], sun[1], sun[2]);
  float lr = len3(rx, ry, rz);
  float snx = sun[0] / ls, sny = sun[1] / ls, snz = sun[2] / ls;
  float rnx = rx / lr, rny = ry / lr, rnz = rz / lr;
  float dt = snx * rnx;
  dt = dt + sny * rny;
  dt = dt + snz * rnz;
  float sv = fminf(fmaxf(dt, 0.0f), 1.2f);
  float p8 = powf(sv, 8.0f), p3 = powf(sv, 3.0f);
  float k = ry * 0.2f;
  r = ((0.6f - k * 1.0f) + 0.075f);
  g = ((0.71f - k * 0.5f) + 0.075f);
  b = ((0.75f - k * 1.0f) + 0.075f);
  r = r + (0.4f * 1.0f) * p8;
  g = g + (0.4f * 0.6f) * p8;
  b = b + (0.4f * 0.1f) * p8;
  r = r + 0.2f * p3;
  g = g + 0.08f * p3;
  b = b + 0.04f * p3;
}

// ---------------------------------------------------------------------------
// dense worlds
// ---------------------------------------------------------------------------
struct DenseWorld {
  const uint8_t* vox;       // material bytes in 4^3 bricks (64 B each), bricks x-fastest
  const uint64_t* l0;       // one bit per voxel of each 4^3 brick
  const uint64_t* l1;       // one bit per brick of each 16^3 region
  const uint32_t* palette;  // RGBA8 colour of material k (the solid atlas cube)
  uint32_t dim;             // voxels per axis (MAP_DIMENSION)
  uint32_t nb;              // bricks per axis (dim / 4)
  uint32_t nr;              // regions per axis (dim / 16)
};

__device__ __forceinline__ size_t dense_byte_index(const DenseWorld& w, int bx, int by, int bz) {
  uint32_t brick = ((uint32_t)(bz >> 2) * w.nb + (uint32_t)(by >> 2)) * w.nb + (uint32_t)(bx >> 2);
  uint32_t local = (uint32_t)(bx & 3) | ((uint32_t)(by & 3) << 2) | ((uint32_t)(bz & 3) << 4);
  return (size_t)brick * 64u + local;
}

__device__ __forceinline__ bool in_dim(const DenseWorld& w, int bx, int by, int bz) {
  return ((uint32_t)bx < w.dim) & ((uint32_t)by < w.dim) & ((uint32_t)bz < w.dim);
}

// Occupancy lookup strategies (identical answers by contract).
struct FlatLookup {
  // returns the material byte (0 = empty); one gather per in-bounds lookup
  __device__ __forceinline__ void reset() {}
  template <bool kStats>
  __device__ __forceinline__ uint32_t lookup(const DenseWorld& w, int bx, int by, int bz,
                                             uint32_t& leaf, uint32_t&, uint32_t&) {
    if (!in_dim(w, bx, by, bz)) return 0;
    if (kStats) leaf++;
    return w.vox[dense_byte_index(w, bx, by, bz)];
  }
  __device__ __forceinline__ uint32_t material(const DenseWorld&, int, int, int, uint32_t v,
                                               uint32_t&) {
    return v;
  }
};

struct BitmaskLookup {
  uint32_t region_key, brick_key;
  uint64_t region_word, brick_word;
  __device__ __forceinline__ void reset() {
    region_key = 0xFFFFFFFFu;
    brick_key = 0xFFFFFFFFu;
    region_word = 0;
    brick_word = 0;
  }
  // returns nonzero when the block is solid; fetches a region word when the
  // 16^3 region changes and a brick word when the 4^3 brick changes and the
  // region word says the brick is not empty.
  template <bool kStats>
  __device__ __forceinline__ uint32_t lookup(const DenseWorld& w, int bx, int by, int bz,
                                             uint32_t&, uint32_t& l1f, uint32_t& l0f) {
    if (!in_dim(w, bx, by, bz)) return 0;
    uint32_t kx = (uint32_t)bx >> 2, ky = (uint32_t)by >> 2, kz = (uint32_t)bz >> 2;
    uint32_t bkey = kx | (ky << 10) | (kz << 20);
    if (bkey != brick_key) {
      brick_key = bkey;
      uint32_t rkey = (kx >> 2) | ((ky >> 2) << 10) | ((kz >> 2) << 20);
      if (rkey != region_key) {
        region_key = rkey;
        uint32_t ridx = ((kz >> 2) * w.nr + (ky >> 2)) * w.nr + (kx >> 2);
        region_word = w.l1[ridx];
        if (kStats) l1f++;
      }
      uint32_t rbit = (kx & 3u) | ((ky & 3u) << 2) | ((kz & 3u) << 4);
      brick_word = 0;
      if ((region_word >> rbit) & 1ull) {
        brick_word = w.l0[(kz * w.nb + ky) * w.nb + kx];
        if (kStats) l0f++;
      }
    }
    uint32_t vbit = (uint32_t)(bx & 3) | ((uint32_t)(by & 3) << 2) | ((uint32_t)(bz & 3) << 4);
    return (uint32_t)(brick_word >> vbit) & 1u;
  }
  __device__ __forceinline__ uint32_t material(const DenseWorld& w, int bx, int by, int bz,
                                               uint32_t, uint32_t& leaf) {
    leaf++;
    return w.vox[dense_byte_index(w, bx, by, bz)];
  }
};

// Result of one traceMap call in a dense world.
struct DenseHit {
  bool hit;
  int face;             // 1..6
  int bx, by, bz;       // block coordinates (pos >> 3)
  uint32_t v;           // lookup value at the hit (flat: the material byte)
  uint32_t sub;         // pos % 8 packed x + 8y + 64z
  float hx, hy, hz;     // hit_pos = gridsCoords + withinGridCoords
  int steps;
};

// Ray state at the block level.
struct DenseRay {
  int Gx, Gy, Gz;     // gridsCoords >> 3
  float wx, wy, wz;   // withinGridCoords, x8 units
  float dx, dy, dz;   // rayDir after the zero substitution
  float ix, iy, iz;   // 1 / rayDir
  float Px, Py, Pz;   // float(rayPositivity << 3)
  float Ax, Ay, Az;   // float((1 - rayPositivity) << 3) * 0.999f
  int Sx, Sy, Sz;     // raySign
  int px, py, pz;     // rayPositivity
  int m;              // minIdx
  int steps, max_steps;
};

// Set-up (map.glsl:85-104) plus iteration 0 (stepSize 0).  Returns true when
// the ray already terminated (hit or miss) in iteration 0; otherwise the ray
// is at the block level, ready for the DDA step of iteration 0.
template <bool kStats, class L>
__device__ __forceinline__ bool dense_begin(DenseRay& r, L& lk, const DenseWorld& w, float ox,
                                            float oy, float oz, float dx, float dy, float dz,
                                            int max_steps, DenseHit& h, uint32_t& leaf,
                                            uint32_t& l1f, uint32_t& l0f) {
  if (dx == 0.0f) dx = 0.001f;
  if (dy == 0.0f) dy = 0.001f;
  if (dz == 0.0f) dz = 0.001f;
  r.dx = dx;
  r.dy = dy;
  r.dz = dz;
  r.Sx = dx > 0.0f ? 1 : (dx < 0.0f ? -1 : 0);
  r.Sy = dy > 0.0f ? 1 : (dy < 0.0f ? -1 : 0);
  r.Sz = dz > 0.0f ? 1 : (dz < 0.0f ? -1 : 0);
  r.px = (1 + r.Sx) >> 1;
  r.py = (1 + r.Sy) >> 1;
  r.pz = (1 + r.Sz) >> 1;
  r.ix = 1.0f / dx;
  r.iy = 1.0f / dy;
  r.iz = 1.0f / dz;
  r.Px = (float)(r.px << 3);
  r.Py = (float)(r.py << 3);
  r.Pz = (float)(r.pz << 3);
  r.Ax = (float)((1 - r.px) << 3) * 0.999f;
  r.Ay = (float)((1 - r.py) << 3) * 0.999f;
  r.Az = (float)((1 - r.pz) << 3) * 0.999f;
  r.m = 0;
  r.steps = 0;
  r.max_steps = max_steps;
  lk.reset();

  float sx = ox * 8.0f, sy = oy * 8.0f, sz = oz * 8.0f;
  int gx = f2i_sat(sx), gy = f2i_sat(sy), gz = f2i_sat(sz);
  float wx = sx - (float)gx, wy = sy - (float)gy, wz = sz - (float)gz;

  h.hit = false;
  h.steps = 0;
  const int B = (int)(8u * w.dim);
  if (max_steps <= 0) return true;
  if (gx >= B || gy >= B || gz >= B || gx < 0 || gy < 0 || gz < 0) return true;
  // iteration 0, stepSize 0: pos = gridsCoords + int(within)
  int qx = gx + (int)wx, qy = gy + (int)wy, qz = gz + (int)wz;
  int bx = qx >> 3, by = qy >> 3, bz = qz >> 3;
  uint32_t v = lk.template lookup<kStats>(w, bx, by, bz, leaf, l1f, l0f);
  if (v != 0u) {
    h.hit = true;
    h.face = 2 - r.px;  // minIdx == 0 on the first iteration
    h.bx = bx;
    h.by = by;
    h.bz = bz;
    h.v = v;
    h.sub = (uint32_t)(qx & 7) | ((uint32_t)(qy & 7) << 3) | ((uint32_t)(qz & 7) << 6);
    h.hx = (float)gx + wx;
    h.hy = (float)gy + wy;
    h.hz = (float)gz + wz;
    h.steps = 1;
    return true;
  }
  // empty block: switch to the block level (map.glsl:140-144)
  int lx = gx & 7, ly = gy & 7, lz = gz & 7;
  r.wx = wx + (float)lx;
  r.wy = wy + (float)ly;
  r.wz = wz + (float)lz;
  r.Gx = (gx - lx) >> 3;
  r.Gy = (gy - ly) >> 3;
  r.Gz = (gz - lz) >> 3;
  return false;
}

// One DDA step (map.glsl:157-162) at the block level followed by the next
// iteration's bounds check and lookup.  Returns true when the ray terminated.
template <bool kStats, class L>
__device__ __forceinline__ bool dense_step(DenseRay& r, L& lk, const DenseWorld& w, DenseHit& h,
                                           uint32_t& leaf, uint32_t& l1f, uint32_t& l0f) {
  float tx = (r.Px - r.wx) * r.ix;
  float ty = (r.Py - r.wy) * r.iy;
  float tz = (r.Pz - r.wz) * r.iz;
  bool xy = tx < ty, xz = tx < tz, yz = ty < tz;
  bool m0 = xy & xz;
  bool m1 = (!xy) & yz;
  bool m2 = !(m0 | m1);
  float tm = m0 ? tx : (m1 ? ty : tz);
  float nwx = r.wx + r.dx * tm;
  float nwy = r.wy + r.dy * tm;
  float nwz = r.wz + r.dz * tm;
  r.wx = m0 ? r.Ax : nwx;
  r.wy = m1 ? r.Ay : nwy;
  r.wz = m2 ? r.Az : nwz;
  r.Gx += m0 ? r.Sx : 0;
  r.Gy += m1 ? r.Sy : 0;
  r.Gz += m2 ? r.Sz : 0;
  r.m = m0 ? 0 : (m1 ? 1 : 2);
  r.steps++;
  if (r.steps >= r.max_steps) {
    h.hit = false;
    h.steps = r.steps;
    return true;
  }
  if (((uint32_t)r.Gx >= w.dim) | ((uint32_t)r.Gy >= w.dim) | ((uint32_t)r.Gz >= w.dim)) {
    h.hit = false;
    h.steps = r.steps;
    return true;
  }
  // pos >> 3 with pos = 8*G + int(within): within lies in (-1, 9)
  int bx = r.Gx + ((int)r.wx >> 3);
  int by = r.Gy + ((int)r.wy >> 3);
  int bz = r.Gz + ((int)r.wz >> 3);
  uint32_t v = lk.template lookup<kStats>(w, bx, by, bz, leaf, l1f, l0f);
  if (v != 0u) {
    h.hit = true;
    h.face = r.m == 0 ? 2 - r.px : (r.m == 1 ? 4 - r.py : 6 - r.pz);
    h.bx = bx;
    h.by = by;
    h.bz = bz;
    h.v = v;
    h.sub = (uint32_t)((int)r.wx & 7) | ((uint32_t)((int)r.wy & 7) << 3) |
            ((uint32_t)((int)r.wz & 7) << 6);
    h.hx = (float)(r.Gx << 3) + r.wx;
    h.hy = (float)(r.Gy << 3) + r.wy;
    h.hz = (float)(r.Gz << 3) + r.wz;
    h.steps = r.steps + 1;
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// reference brickmap + atlas layout (literal two-level traceMap)
// ---------------------------------------------------------------------------
struct BrickmapWorld {
  const uint32_t* chunks;  // (dim/8)^3, 0 = empty else pool index + 1
  const uint32_t* pool;    // 512 u32 per block, in-block index x + 8y + 64z
  const uint32_t* atlas;   // 256^3 RGBA8 texels
  uint32_t dim;            // MAP_DIMENSION
};

// map_getVoxel (map.glsl:38-47) of block coordinates b
__device__ __forceinline__ uint32_t bm_block(const BrickmapWorld& w, int bx, int by, int bz) {
  int M = (int)(w.dim >> 3);
  int cx = bx >> 3, cy = by >> 3, cz = bz >> 3;
  if (cx < 0 || cy < 0 || cz < 0 || cx >= M || cy >= M || cz >= M) return 0;
  uint32_t idx = w.chunks[cx + M * (cy + cz * M)];
  if (idx == 0) return 0;
  size_t o = (size_t)(idx - 1u) * 512u + (size_t)(bx % 8) + (size_t)(((bz % 8) * 8 + (by % 8)) * 8);
  return w.pool[o];
}

// map_getSubVoxel (map.glsl:57-60)
__device__ __forceinline__ uint32_t bm_sub(const BrickmapWorld& w, uint32_t block, int sx, int sy,
                                           int sz) {
  // sx, sy, sz = pos % 8 lie in [0, 8) (pos >= 0 once the bounds check
  // passed); the & 255 only keeps the texel inside the 256^3 atlas by
  // construction.
  uint32_t id = block & 0xFFFFFFFu;
  uint32_t x = ((uint32_t)sx + (id & 31u) * 8u) & 255u;
  uint32_t y = ((uint32_t)sy + ((id / 32u) & 31u) * 8u) & 255u;
  uint32_t z = ((uint32_t)sz + ((id / 1024u) & 31u) * 8u) & 255u;
  return w.atlas[x + 256u * (y + 256u * z)];
}

struct GeneralHit {
  uint32_t data;
  int face;
  int bx, by, bz;
  uint32_t sub;
  float hx, hy, hz;
  int steps;
};

// traceMap (map.glsl:83-168) literally, both levels.
template <bool kStats>
__device__ bool trace_general(const BrickmapWorld& w, float ox, float oy, float oz, float dx,
                              float dy, float dz, int max_steps, GeneralHit& h, uint32_t& leaf) {
  if (dx == 0.0f) dx = 0.001f;
  if (dy == 0.0f) dy = 0.001f;
  if (dz == 0.0f) dz = 0.001f;
  const int B = (int)(8u * w.dim);
  int sgn[3] = {dx > 0.0f ? 1 : (dx < 0.0f ? -1 : 0), dy > 0.0f ? 1 : (dy < 0.0f ? -1 : 0),
                dz > 0.0f ? 1 : (dz < 0.0f ? -1 : 0)};
  int pv[3] = {(1 + sgn[0]) >> 1, (1 + sgn[1]) >> 1, (1 + sgn[2]) >> 1};
  float d[3] = {dx, dy, dz};
  float inv[3] = {1.0f / dx, 1.0f / dy, 1.0f / dz};
  int m = 0;
  float o8[3] = {ox * 8.0f, oy * 8.0f, oz * 8.0f};
  int g[3] = {f2i_sat(o8[0]), f2i_sat(o8[1]), f2i_sat(o8[2])};
  float wc[3] = {o8[0] - (float)g[0], o8[1] - (float)g[1], o8[2] - (float)g[2]};
  int s = 0;
  h.data = 0;
  h.hx = h.hy = h.hz = -1.0f;
  int step;
  for (step = 0; step < max_steps; step++) {
    if (g[0] >= B || g[1] >= B || g[2] >= B || g[0] < 0 || g[1] < 0 || g[2] < 0) break;
    int p[3] = {g[0] + (int)wc[0], g[1] + (int)wc[1], g[2] + (int)wc[2]};
    int b[3] = {p[0] >> 3, p[1] >> 3, p[2] >> 3};
    uint32_t block = bm_block(w, b[0], b[1], b[2]);
    if (kStats) leaf++;
    if (block != 0u) {
      uint32_t sub = bm_sub(w, block, p[0] % 8, p[1] % 8, p[2] % 8);
      if (kStats) leaf++;
      if (sub != 0u) {
        h.data = sub;
        h.face = m == 0 ? 2 - pv[0] : (m == 1 ? 4 - pv[1] : 6 - pv[2]);
        h.bx = b[0];
        h.by = b[1];
        h.bz = b[2];
        h.sub = (uint32_t)(p[0] & 7) | ((uint32_t)(p[1] & 7) << 3) | ((uint32_t)(p[2] & 7) << 6);
        h.hx = (float)g[0] + wc[0];
        h.hy = (float)g[1] + wc[1];
        h.hz = (float)g[2] + wc[2];
        h.steps = step + 1;
        return true;
      }
      if (s != 0) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
          g[i] += (int)wc[i];
          wc[i] = wc[i] - floorf(wc[i]);
        }
        s = 0;
      }
    } else if (s != 3) {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        int lo = g[i] & 7;
        wc[i] = wc[i] + (float)lo;
        g[i] -= lo;
      }
      s = 3;
    }
    float t[3];
#pragma unroll
    for (int i = 0; i < 3; i++) t[i] = ((float)(pv[i] << s) - wc[i]) * inv[i];
    m = t[0] < t[1] ? (t[0] < t[2] ? 0 : 2) : (t[1] < t[2] ? 1 : 2);
    float tm = m == 0 ? t[0] : (m == 1 ? t[1] : t[2]);
    int sm = m == 0 ? sgn[0] : (m == 1 ? sgn[1] : sgn[2]);
    int pm = m == 0 ? pv[0] : (m == 1 ? pv[1] : pv[2]);
    float am = (float)((1 - pm) << s) * 0.999f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      float nw = wc[i] + d[i] * tm;
      wc[i] = (i == m) ? am : nw;
      g[i] += (i == m) ? sm * (1 << s) : 0;
    }
  }
  h.steps = step;
  return false;
}

}  // namespace vt
