// vt_trace.hpp -- device-side REF voxel traversal for gfx950.
//
// Restates traceMap (assets/shaders/map.glsl:83-168) and its callers
// (primary.comp.glsl:23-69, secondary.comp.glsl:18-50, blit.fragment.glsl:30-31)
// in binary32 with the same operation order as the CPU restatement: built with
// -ffp-contract=off and correctly rounded division/sqrt, every multiply and add
// rounds once, so hits are bit-exact.
//
// Two traversal loops:
//  * dense_begin + dense_run: worlds where one grid voxel is one reference
//    block whose atlas model is a solid cube.  After the first lookup
//    traceMap never leaves the block level (stepSize 3), so the block
//    coordinates G = gridsCoords >> 3 live in a Morton code while
//    withinGridCoords stays in x8 units; this is the same recurrence, with
//    pos >> 3 == G + (int(within) >> 3).
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

// RGBA8 framebuffer store of a colour: per channel clamp to [0,1], round to
// nearest (the unorm8 rule), R in the low byte.
__device__ __forceinline__ uint32_t pack_unorm8(float4 c) {
  const float v[4] = {c.x, c.y, c.z, c.w};
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float k = fminf(fmaxf(v[i], 0.0f), 1.0f);
    r |= (uint32_t)floorf(k * 255.0f + 0.5f) << (8 * i);
  }
  return r;
}

// Present pass of one pixel (blit.fragment.glsl:32-35): texPos is the pixel
// centre of the full-screen quad, ((x+0.5)/W, (y+0.5)/H).
__device__ __forceinline__ float4 present_pixel(float4 c, uint32_t x, uint32_t y, uint32_t W,
                                                uint32_t H) {
  const float u = ((float)x + 0.5f) / (float)W;
  const float v = ((float)y + 0.5f) / (float)H;
  const float du = u - 0.5f, dv = v - 0.5f;
  float d2 = du * du;
  d2 = d2 + dv * dv;
  if (sqrtf(d2) <= 0.002f) {  // crosshair: mix(color, vec4(1,1,1,0.4), 0.5)
    c = make_float4(c.x * 0.5f + 0.5f, c.y * 0.5f + 0.5f, c.z * 0.5f + 0.5f,
                    c.w * 0.5f + 0.4f * 0.5f);
  }
  const float a = u * (1.0f - u), b = v * (1.0f - v);
  const float g = powf((a * b) * 15.0f, 0.6f * 0.3f);  // vignette
  return make_float4(g * c.x, g * c.y, g * c.z, g * c.w);
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
// dense worlds: Morton-ordered voxels
// ---------------------------------------------------------------------------
// The material bytes are stored in Morton (Z) order: voxel (x,y,z) at
// code = interleave(x, y, z) (x in bits 0,3,6,..., y in 1,4,..., z in 2,5,...).
// A 128-byte cache line then holds a ~5^3 neighbourhood, a 4^3 brick is 64
// consecutive codes and a 16^3 region 4096, whatever direction a ray moves.
// The two occupancy levels fall out of the same order: L0 word code >> 6 has
// bit code & 63, L1 word code >> 12 has bit (code >> 6) & 63.
struct DenseWorld {
  const uint8_t* vox;       // material byte of voxel (x,y,z) at morton(x,y,z)
  const uint64_t* l0;       // 4^3 bricks: bit per voxel
  const uint64_t* l1;       // 16^3 regions: bit per nonempty 4^3 brick
  const uint64_t* l1b;      // 32^3 regions: bit per nonempty 8^3 brick
  const uint32_t* palette;  // RGBA8 colour of material k (the solid atlas cube)
  uint32_t dim;             // voxels per axis (MAP_DIMENSION), a power of two
  uint32_t log2dim;
  uint32_t mx, my, mz;      // dilated masks of x, y, z over 3*log2(dim) bits
  uint32_t nvox;            // dim^3 (bytes of vox)
  const uint32_t* prefix;   // compact albedo index: occupied voxels before each
                            // L0 word (null until built)
  // free-region map (see FreeMap below); free_bytes 0 = rays are not ended early
  const uint32_t* freemap;
  uint32_t free_bytes, free_oct_bytes;
  // the box of the nonempty voxels widened by the traversal's reach
  // (line_meets_reach); the whole grid when the world is empty; everywhere
  // (reach_everywhere) when step counts are observed
  float reach_lo[3], reach_hi[3];
};

__device__ __forceinline__ uint32_t part1by2(uint32_t x) {
  x &= 0x000003FFu;
  x = (x | (x << 16)) & 0xFF0000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__device__ __forceinline__ uint32_t compact1by2(uint32_t x) {
  x &= 0x09249249u;
  x = (x | (x >> 2)) & 0x030C30C3u;
  x = (x | (x >> 4)) & 0x0300F00Fu;
  x = (x | (x >> 8)) & 0xFF0000FFu;
  x = (x | (x >> 16)) & 0x000003FFu;
  return x;
}
__device__ __forceinline__ uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}

__device__ __forceinline__ bool in_dim(const DenseWorld& w, int bx, int by, int bz) {
  return ((uint32_t)bx < w.dim) & ((uint32_t)by < w.dim) & ((uint32_t)bz < w.dim);
}

// Steps per iteration of the empty-space phase (dense_skip_empty): 6 for
// frames with a step cap of at least 512, 5 below.  Measured on MI355X with
// the free-region test: C1 at 1080p 0.2740 / 0.2725 / 0.2710 ms at K = 6 /
// 7 / 8 against 0.2796 at 5; C3's views 0.1065 ms against 0.1117; the C1
// world in the reference layout 0.2858 (6) / 0.2892 (7) / 0.292 (8) against
// 0.2869 (5); the engine's own frame (caps 192/48, short runs between model
// blocks) 0.2747 at 5 against 0.2972 at 6 and 0.335 at 7.
constexpr int kSkipShort = 5;
constexpr int kSkipLong = 6;
constexpr int kLongCap = 512;

// Occupancy probe strategies (identical answers by contract).
struct FlatLookup {
  // one byte gather per probe; returns the material (0 = empty)
  static constexpr bool kPrefetch = true;  // probes have no state: run ahead
  static constexpr bool kFree = false;     // no free-region test
  static constexpr int kSkip = kSkipShort;  // empty-space iteration length
  __device__ __forceinline__ void reset() {}
  template <bool kStats>
  __device__ __forceinline__ uint32_t probe(const DenseWorld& w, uint32_t code, uint32_t&,
                                            uint32_t&) {
    return w.vox[code];
  }
  __device__ __forceinline__ uint32_t material(const DenseWorld&, uint32_t, uint32_t v,
                                               uint32_t&) {
    return v;
  }
};

// The 2-level occupancy probe: a region word (one bit per brick) and the
// probed voxel's 64-bit leaf word, both kept in registers between probes.
//  kBrick = 4: 4^3 bricks in 16^3 regions -- region word l1[code >> 12],
//              bit (code >> 6) & 63; the brick is the leaf word l0[code >> 6].
//  kBrick = 8: 8^3 bricks in 32^3 regions -- region word l1b[code >> 15],
//              bit (code >> 9) & 63; an occupied 8^3 brick's voxels are the
//              8 leaf words l0[code >> 6], of which the one holding the
//              probed 4^3 sub-brick is fetched.
// A leaf word is fetched only when its brick is marked occupied, a region
// word only when the region changes.
template <uint32_t kBrick>
struct BitmaskLookupT {
  static_assert(kBrick == 4 || kBrick == 8, "4^3 or 8^3 bricks");
  static constexpr bool kPrefetch = false;
  static constexpr bool kFree = false;
  static constexpr int kSkip = kSkipShort;  // (the plain loop has no empty-space phase)
  static constexpr uint32_t kBrickShift = kBrick == 8 ? 9u : 6u;  // log2 codes per brick
  uint32_t brick, region, leaf;
  uint64_t bword, rword;
  bool bset;  // the current brick is marked occupied
  __device__ __forceinline__ void reset() {
    brick = 0xFFFFFFFFu;
    region = 0xFFFFFFFFu;
    leaf = 0xFFFFFFFFu;
    bword = 0;
    rword = 0;
    bset = false;
  }
  template <bool kStats>
  __device__ __forceinline__ uint32_t probe(const DenseWorld& w, uint32_t code, uint32_t& l1f,
                                            uint32_t& l0f) {
    const uint32_t lf = code >> 6;
    if (lf != leaf) {
      leaf = lf;
      const uint32_t bk = code >> kBrickShift;
      if (bk != brick) {
        brick = bk;
        const uint32_t rk = bk >> 6;
        if (rk != region) {
          region = rk;
          rword = kBrick == 8 ? w.l1b[rk] : w.l1[rk];
          if (kStats) l1f++;
        }
        bset = ((rword >> (bk & 63u)) & 1ull) != 0ull;
      }
      bword = 0;
      if (bset) {
        bword = w.l0[lf];
        if (kStats) l0f++;
      }
    }
    return (uint32_t)(bword >> (code & 63u)) & 1u;
  }
  __device__ __forceinline__ uint32_t material(const DenseWorld& w, uint32_t code, uint32_t,
                                               uint32_t& leaf_reads) {
    leaf_reads++;
    return w.vox[code];
  }
};
using BitmaskLookup = BitmaskLookupT<4>;
using BitmaskLookup8 = BitmaskLookupT<8>;

// The flat lookup whose empty-space loops also test the free-region map (see
// free_rsrc below) and end a ray that can hit nothing more.
struct FlatLookupFree : FlatLookup {
  static constexpr bool kFree = true;
  static constexpr int kSkip = kSkipLong;
};

// The flat lookup for frames with long step caps: the longer empty-space
// iteration, no free-region test.
struct FlatLookupLong : FlatLookup {
  static constexpr int kSkip = kSkipLong;
};

// Result of one traceMap call in a dense world.
struct DenseHit {
  bool hit;
  int face;             // 1..6
  int bx, by, bz;       // block coordinates (pos >> 3)
  uint32_t code;        // morton(bx, by, bz)
  uint32_t v;           // probe value at the hit (flat: the material byte)
  uint32_t sub;         // pos % 8 packed x + 8y + 64z
  float hx, hy, hz;     // hit_pos = gridsCoords + withinGridCoords
  int steps;
};

// Ray state at the block level.
struct DenseRay {
  float wx, wy, wz;   // withinGridCoords, x8 units
  float dx, dy, dz;   // rayDir after the zero substitution
  float ix, iy, iz;   // 1 / rayDir
  float Px, Py, Pz;   // float(rayPositivity << 3)
  float Ax, Ay, Az;   // float((1 - rayPositivity) << 3) * 0.999f
  int px, py, pz;     // rayPositivity
  uint32_t c;         // Morton code of G = gridsCoords >> 3, axes the ray
                      // travels negatively stored complemented
  uint32_t flip;      // the complemented axes' masks; address = c ^ flip
  int steps, max_steps;
  int blind;          // steps whose lookups can only read empty (blind_steps);
                      // after dense_run_prefetch, the steps taken without them
};

// The per-ray constants of traceMap's set-up (map.glsl:85-97): the 0.001
// substitution of zero direction components, positivity, 1/d, and the
// block-level crossing and re-anchoring constants.  A __host__ __device__
// function: the shadow rays' direction is the sun for every pixel, so their
// constants are computed once per launch on the host (binary32, correctly
// rounded division, like the device) and reach the kernels as arguments --
// scalar registers instead of per-lane copies of uniform values.
struct RayConsts {
  float d[3];    // rayDir after the zero substitution
  float inv[3];  // 1 / rayDir
  float P[3];    // float(rayPositivity << 3)
  float A[3];    // float((1 - rayPositivity) << 3) * 0.999f
  int pos[3];    // rayPositivity
  uint32_t flip; // masks of the axes the ray travels negatively
};

__host__ __device__ inline RayConsts make_ray_consts(float dx, float dy, float dz, uint32_t mx,
                                                     uint32_t my, uint32_t mz) {
  RayConsts k;
  const float d[3] = {dx == 0.0f ? 0.001f : dx, dy == 0.0f ? 0.001f : dy,
                      dz == 0.0f ? 0.001f : dz};
  const uint32_t m[3] = {mx, my, mz};
  k.flip = 0u;
  for (int i = 0; i < 3; i++) {
    const int sg = d[i] > 0.0f ? 1 : (d[i] < 0.0f ? -1 : 0);
    k.d[i] = d[i];
    k.pos[i] = (1 + sg) >> 1;
    k.inv[i] = 1.0f / d[i];
    k.P[i] = (float)(k.pos[i] << 3);
    k.A[i] = (float)((1 - k.pos[i]) << 3) * 0.999f;
    k.flip |= k.pos[i] ? 0u : m[i];
  }
  return k;
}

__device__ __forceinline__ void dense_ray_from(DenseRay& r, const RayConsts& k) {
  r.dx = k.d[0];
  r.dy = k.d[1];
  r.dz = k.d[2];
  r.ix = k.inv[0];
  r.iy = k.inv[1];
  r.iz = k.inv[2];
  r.Px = k.P[0];
  r.Py = k.P[1];
  r.Pz = k.P[2];
  r.Ax = k.A[0];
  r.Ay = k.A[1];
  r.Az = k.A[2];
  r.px = k.pos[0];
  r.py = k.pos[1];
  r.pz = k.pos[2];
  r.flip = k.flip;
}

__device__ __forceinline__ void dense_ray_consts(DenseRay& r, const DenseWorld& w, float dx,
                                                 float dy, float dz) {
  dense_ray_from(r, make_ray_consts(dx, dy, dz, w.mx, w.my, w.mz));
}

// Can traceMap's block-level walk from e (voxel units) along d probe a
// nonempty voxel?  Only if the line e + d t, t >= 0, meets the box of the
// nonempty voxels widened by how far a probe can lie from that line:
//   * every step moves the position along d up to binary32 rounding (< 1e-6
//     sub-voxel per step: 0.01 voxel over 65,535 steps);
//   * a step across a negative axis re-anchors within at 8 x 0.999 instead
//     of 8, 0.001 voxel off the line per crossing, at most dim crossings per
//     axis inside the grid (the walk ends where it leaves it);
//   * the probed voxel is the position's, or the next one on a positive axis
//     whose within reached 8: within 2 voxels of the position;
// so the widening of 2.5 + dim / 1000 voxels (wide_margin) covers all three
// with room for the rounding of this test.  A ray whose line misses it ends
// as a miss; only its step count differs from traceMap's.
__host__ __device__ __forceinline__ float wide_margin(uint32_t dim) {
  return 2.5f + (float)dim / 1000.0f;
}
// A reach every line meets: for launches whose step counts are observed
// (stored per pixel, or statistics of traceMap's steps), where no ray may
// end before traceMap does.
__host__ __device__ inline void reach_everywhere(float* lo, float* hi) {
  for (int i = 0; i < 3; i++) {
    lo[i] = -1.0e30f;
    hi[i] = 1.0e30f;
  }
}
// The slab test of the line against [lo, hi] (inv = 1 / d after the zero
// substitution, so never infinite).
__device__ __forceinline__ bool line_meets_box(const float* lo, const float* hi, float ex,
                                               float ey, float ez, const float* inv) {
  const float e[3] = {ex, ey, ez};
  float tn = 0.0f, tf = 3.0e38f;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float a = (lo[i] - e[i]) * inv[i], b = (hi[i] - e[i]) * inv[i];
    tn = fmaxf(tn, fminf(a, b));
    tf = fminf(tf, fmaxf(a, b));
  }
  return tn <= tf;
}
__device__ __forceinline__ bool line_meets_reach(const DenseWorld& w, float ex, float ey,
                                                 float ez, const RayConsts& k) {
  return line_meets_box(w.reach_lo, w.reach_hi, ex, ey, ez, k.inv);
}

// How many of traceMap's steps from e (voxel units) along k.d, whose line
// meets the reach box, come before the walk can probe a nonempty voxel: the
// blocks the line crosses before it enters the reach box, less a slack.
// Why the lookups of the first blind_steps steps read empty: after k steps
// the walk has crossed k block boundaries, c_i of them on axis i, and its
// position lies within the drift (< 1.03 voxel, see above) of a point
// L(s) of the line; the boundaries the line crosses up to s, n_i(s), differ
// from c_i by at most 4 (2 for the drift, 1 for a within that reached 8, 1
// for the binary32 floor below), so k <= n(t_enter) - 13 gives n(s) <
// n(t_enter), s < t_enter, L(s) outside the reach box, and the probe
// outside the box of the nonempty voxels (the reach box's argument).
__device__ __forceinline__ int blind_steps_box(const float* lo, const float* hi, float ex,
                                               float ey, float ez, const float* d,
                                               const float* inv) {
  const float e[3] = {ex, ey, ez};
  float tn = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const float a = (lo[i] - e[i]) * inv[i], b = (hi[i] - e[i]) * inv[i];
    tn = fmaxf(tn, fminf(a, b));
  }
  int n = 0;
#pragma unroll
  for (int i = 0; i < 3; i++) n += abs((int)floorf(e[i] + d[i] * tn) - (int)floorf(e[i]));
  return n - 13;
}
__device__ __forceinline__ int blind_steps(const DenseWorld& w, float ex, float ey, float ez,
                                           const RayConsts& k) {
  return blind_steps_box(w.reach_lo, w.reach_hi, ex, ey, ez, k.d, k.inv);
}

// Set-up (map.glsl:85-104) plus iteration 0 (stepSize 0).  Returns true when
// the ray already terminated (hit or miss) in iteration 0; otherwise the ray
// is at the block level, ready for the DDA step of iteration 0.
// dense_begin_k: the same with the direction's constants made beforehand.
template <bool kStats, class L>
__device__ __forceinline__ bool dense_begin_k(DenseRay& r, L& lk, const DenseWorld& w, float ox,
                                              float oy, float oz, const RayConsts& k,
                                              int max_steps, DenseHit& h, uint32_t& l1f,
                                              uint32_t& l0f);

template <bool kStats, class L>
__device__ __forceinline__ bool dense_begin(DenseRay& r, L& lk, const DenseWorld& w, float ox,
                                            float oy, float oz, float dx, float dy, float dz,
                                            int max_steps, DenseHit& h, uint32_t& l1f,
                                            uint32_t& l0f) {
  return dense_begin_k<kStats>(r, lk, w, ox, oy, oz, make_ray_consts(dx, dy, dz, w.mx, w.my, w.mz),
                               max_steps, h, l1f, l0f);
}

template <bool kStats, class L>
__device__ __forceinline__ bool dense_begin_k(DenseRay& r, L& lk, const DenseWorld& w, float ox,
                                              float oy, float oz, const RayConsts& k,
                                              int max_steps, DenseHit& h, uint32_t& l1f,
                                              uint32_t& l0f) {
  dense_ray_from(r, k);
  r.steps = 0;
  r.max_steps = max_steps;
  r.blind = 0;
  lk.reset();

  float s8x = ox * 8.0f, s8y = oy * 8.0f, s8z = oz * 8.0f;
  int gx = f2i_sat(s8x), gy = f2i_sat(s8y), gz = f2i_sat(s8z);
  float wx = s8x - (float)gx, wy = s8y - (float)gy, wz = s8z - (float)gz;

  h.hit = false;
  h.steps = 0;
  const int B = (int)(8u * w.dim);
  if (max_steps <= 0) return true;
  if (gx >= B || gy >= B || gz >= B || gx < 0 || gy < 0 || gz < 0) return true;
  // iteration 0, stepSize 0: pos = gridsCoords + int(within)
  int qx = gx + (int)wx, qy = gy + (int)wy, qz = gz + (int)wz;
  int bx = qx >> 3, by = qy >> 3, bz = qz >> 3;
  uint32_t v = 0;
  uint32_t code = 0;
  if (in_dim(w, bx, by, bz)) {
    code = morton3((uint32_t)bx, (uint32_t)by, (uint32_t)bz);
    v = lk.template probe<kStats>(w, code, l1f, l0f);
  }
  if (v != 0u) {
    h.hit = true;
    h.face = 2 - r.px;  // minIdx == 0 on the first iteration
    h.bx = bx;
    h.by = by;
    h.bz = bz;
    h.code = code;
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
  r.c = morton3((uint32_t)(gx - lx) >> 3, (uint32_t)(gy - ly) >> 3, (uint32_t)(gz - lz) >> 3) ^
        r.flip;
  return false;
}

// Real-space increment of one axis of a Morton code (mask M); sets `out`
// when the coordinate was already dim - 1.
__device__ __forceinline__ uint32_t morton_inc(uint32_t a, uint32_t M, bool& out) {
  out |= (a & M) == M;
  return (((a | ~M) + 1u) & M) | (a & ~M);
}

// The block-level loop of traceMap (map.glsl:106-165 at stepSize 3) after
// iteration 0.  A step is the DDA arithmetic (map.glsl:157-162), one dilated
// increment of the crossed axis in the flipped Morton code and one probe:
//   M = mask of the crossed axis,  c' = ((c & M) - M) & M | (c & ~M)
// Leaving the grid is exactly the wrap of that increment, (c & M) == M,
// which is traceMap's bounds test.  The probed block is pos >> 3 =
// G + (int(within) >> 3): within reaches 8 on a positive axis the DDA did not
// cross when two crossing times tie exactly (frequent for rays started on
// lattice points, e.g. every shadow ray), so that correction is a short
// branch inside the loop rather than an exit from it.
// A wave-uniform value held in a VGPR for the length of a loop.  The step's
// mask select (two selects between the three axis masks) cannot read two
// SGPRs in one instruction, and the compiler re-creates VGPR copies of the
// masks with a v_mov at every step rather than keep three registers; an
// opaque copy cannot be re-created, so it is made once and stays.
__device__ __forceinline__ uint32_t in_vgpr(uint32_t s) {
  uint32_t v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// One block-level DDA step of traceMap (map.glsl:157-162) and the cell it
// lands in; the state after the step is what the next lookup probes.
struct DenseCell {
  float wx, wy, wz;  // withinGridCoords
  uint32_t c;        // flipped Morton code of gridsCoords >> 3
  uint32_t M;        // mask of the axis crossed by the step (minIdx)
  uint32_t a;        // Morton code of the probed block pos >> 3, or
                     // kOutside when the within >= 8 correction left the grid
  bool wrapped;      // the step left the grid (traceMap's bounds test)
};

// Beyond every grid (codes use at most 30 bits): a bounds-checked buffer load
// of it returns 0, i.e. "empty", as map_getChunkFlags does outside the map.
constexpr uint32_t kOutside = 0xFFFFFFFFu;

// Real-space increment of one axis of a Morton code; kOutside when the
// coordinate was already dim - 1 (kOutside stays kOutside).
__device__ __forceinline__ uint32_t morton_inc_or_out(uint32_t a, uint32_t M) {
  return (a & M) == M ? kOutside : ((((a | ~M) + 1u) & M) | (a & ~M));
}

// The step's ray constants are passed as scalars: the mask select below must
// stay a select of values (a struct here was turned into a dynamically
// indexed stack object, one scratch load per step).
__device__ __forceinline__ DenseCell dense_step(float Px, float Py, float Pz, float Ax, float Ay,
                                                float Az, float dx, float dy, float dz, float ix,
                                                float iy, float iz, uint32_t mx, uint32_t my,
                                                uint32_t mz, uint32_t flip, const DenseCell& s) {
  DenseCell n;
  const float tx = (Px - s.wx) * ix;
  const float ty = (Py - s.wy) * iy;
  const float tz = (Pz - s.wz) * iz;
  // minIdx = tx < ty ? (tx < tz ? 0 : 2) : (ty < tz ? 1 : 2): m1 = yz & !m0
  // is the same rule.  t is never NaN (d != 0 after the 0.001 substitution,
  // within is finite), so t[minIdx] is the smallest t; min3 can differ from
  // it only in the sign of a zero, which no later operation can see
  // (w + d * (+-0) = w, P - w, int(w) and hit_pos are unchanged).
  const bool xy = tx < ty, xz = tx < tz, yz = ty < tz;
  const bool m0 = xy & xz;
  const bool m1 = yz & !m0;
  const float tm = fminf(fminf(tx, ty), tz);
  const float nwx = s.wx + dx * tm;
  const float nwy = s.wy + dy * tm;
  const float nwz = s.wz + dz * tm;
  n.wx = m0 ? Ax : nwx;
  n.wy = m1 ? Ay : nwy;
  n.wz = (m0 | m1) ? nwz : Az;
  n.M = m0 ? mx : (m1 ? my : mz);
  const uint32_t cm = s.c & n.M;
  n.wrapped = cm == n.M;
  n.c = ((cm - n.M) & n.M) | (s.c & ~n.M);
  n.a = n.c ^ flip;
  if (fmaxf(fmaxf(n.wx, n.wy), n.wz) >= 8.0f) {  // pos >> 3 = G + (int(within) >> 3)
    if (n.wx >= 8.0f) n.a = morton_inc_or_out(n.a, mx);
    if (n.wy >= 8.0f) n.a = morton_inc_or_out(n.a, my);
    if (n.wz >= 8.0f) n.a = morton_inc_or_out(n.a, mz);
  }
  return n;
}

// The flat lookup's loop, software-pipelined: the byte of step k+1's block is
// requested before step k's byte is tested, so each gather's latency overlaps
// one step of DDA arithmetic; the run-ahead result is discarded when the ray
// stops at step k.  The bytes are read with bounds-checked buffer loads
// (base and size in a scalar resource, a 32-bit offset per lane): a wrapped
// code is still a code of the grid, and kOutside reads as 0.  Same steps,
// same stopping rule and same final state as the plain loop.
// Block-level steps from s0 (whose own lookup was empty, `steps0` steps
// done) over the byte array behind `bytes` until a step lands on a nonzero
// byte, leaves the grid or reaches max_steps.  Returns the stopping state
// field by field (a struct select would go through memory).
__device__ __forceinline__ void dense_loop_prefetch(
    float Px, float Py, float Pz, float Ax, float Ay, float Az, float dx, float dy, float dz,
    float ix, float iy, float iz, uint32_t mx, uint32_t my, uint32_t mz, uint32_t flip,
    const DenseCell& s0, int steps0, int max_steps, __amdgpu_buffer_rsrc_t bytes, float& fwx,
    float& fwy, float& fwz, uint32_t& fc, uint32_t& fM, uint32_t& fa, uint32_t& fv, int& sF,
    bool& wrappedF) {
#define VT_STEP(s) dense_step(Px, Py, Pz, Ax, Ay, Az, dx, dy, dz, ix, iy, iz, mx, my, mz, flip, s)
#define VT_LOAD(a) (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(bytes, (int)(a), 0, 0)
  DenseCell A = VT_STEP(s0), B;
  int sA = steps0 + 1, sB = 0;
  uint32_t vA = VT_LOAD(A.a), vB = 0u;
  bool useB;
  for (;;) {
    B = VT_STEP(A);
    vB = VT_LOAD(B.a);
    if (A.wrapped | (vA != 0u) | (sA >= max_steps)) {
      useB = false;
      break;
    }
    sB = sA + 1;
    A = VT_STEP(B);
    vA = VT_LOAD(A.a);
    if (B.wrapped | (vB != 0u) | (sB >= max_steps)) {
      useB = true;
      break;
    }
    sA = sB + 1;
  }
#undef VT_STEP
#undef VT_LOAD
  sF = useB ? sB : sA;
  wrappedF = useB ? B.wrapped : A.wrapped;
  fwx = useB ? B.wx : A.wx;
  fwy = useB ? B.wy : A.wy;
  fwz = useB ? B.wz : A.wz;
  fc = useB ? B.c : A.c;
  fM = useB ? B.M : A.M;
  fa = useB ? B.a : A.a;
  fv = useB ? vB : vA;
}

// How an empty-space run ended.
enum SkipStatus : int {
  kSkipLimit = 0,    // reached its step limit; every lookup so far empty
  kSkipBlocked = 1,  // an iteration met a nonzero byte or left the grid
  kSkipFree = 2,     // the ray is in a free region: it can hit nothing more
};

// The free-region map (built by launch_build_freemap): regions of 16^3
// blocks, region code = block Morton code >> 12; bit r of the bitmap of
// direction octant o is set when no region at or ahead of r on every axis in
// o's travel direction (r included) holds a nonempty block.
//
// Why a ray standing in such a region can hit nothing more, so that ending
// it as a miss changes no output but its step count: a step of traceMap's
// block-level recurrence moves the crossed axis one block in the travel
// direction and no axis the other way, and a lookup probes the state's block
// G or, on a positive axis whose within reached 8, the block after it.  So
// every later lookup is a block at or ahead of G on every axis -- inside the
// forward orthant of G's region, or outside the grid, which reads as empty --
// whatever the rounding and the 0.999 re-anchoring do to within.  With that
// orthant empty every later lookup reads 0, and traceMap ends by leaving the
// grid or at its step cap: a miss either way.  Octant bit k = the ray
// travels negatively on axis k; free_bytes 0 reads as "never free" through
// the bounds-checked load.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t free_rsrc(const uint32_t* fm, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)fm, (short)0, (int)bytes, 0x00020000);
}

// Byte offset of the octant bitmap of a ray whose complemented (negative)
// axes are `flip`.
__device__ __forceinline__ uint32_t free_octant_offset(uint32_t flip, uint32_t mx, uint32_t my,
                                                       uint32_t mz, uint32_t oct_bytes) {
  const uint32_t o = ((flip & mx) ? 1u : 0u) | ((flip & my) ? 2u : 0u) | ((flip & mz) ? 4u : 0u);
  return o * oct_bytes;
}

// The free-map word holding region (block code >> 12) of real block code g.
__device__ __forceinline__ uint32_t free_word(__amdgpu_buffer_rsrc_t fm, uint32_t fofs,
                                             uint32_t g) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(fm, (int)(fofs + ((g >> 17) << 2)), 0, 0);
}
__device__ __forceinline__ bool free_bit(uint32_t word, uint32_t g) {
  return ((word >> ((g >> 12) & 31u)) & 1u) != 0u;
}

// Steps K at a time from s while steps + K <= lim without their lookups:
// the caller guarantees that they all read empty (blind_steps).  Stops
// before an iteration that leaves the grid; the state is traceMap's after
// `steps` steps, for the empty-space phase to continue from.
template <int K>
__device__ __forceinline__ void dense_skip_blind(float Px, float Py, float Pz, float Ax,
                                                 float Ay, float Az, float dx, float dy, float dz,
                                                 float ix, float iy, float iz, uint32_t mx,
                                                 uint32_t my, uint32_t mz, uint32_t flip,
                                                 DenseCell& s, int& steps, int lim) {
  float bx = s.wx, by = s.wy, bz = s.wz;
  uint32_t bc = s.c;
  int sb = steps;
  while (sb + K <= lim) {
    DenseCell n;
    n.wx = bx;
    n.wy = by;
    n.wz = bz;
    n.c = bc;
    bool wrapped = false;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      n = dense_step(Px, Py, Pz, Ax, Ay, Az, dx, dy, dz, ix, iy, iz, mx, my, mz, flip, n);
      wrapped |= n.wrapped;
    }
    if (wrapped) break;
    bx = n.wx;
    by = n.wy;
    bz = n.wz;
    bc = n.c;
    sb += K;
  }
  s.wx = bx;
  s.wy = by;
  s.wz = bz;
  s.c = bc;
  steps = sb;
}

// Empty-space phase of the flat loop.  An iteration takes K block-level steps
// from `s` (whose own lookup was empty, `steps` steps done) and requests all K
// bytes without testing any of them in between; only when all K are zero and
// no step left the grid does the state advance by K steps.  Otherwise `s` and
// `steps` are left at the start of that iteration and the exact loop
// (dense_loop_prefetch) replays it to the first stopping step.  The state
// after K empty, in-grid steps is exactly the exact loop's state after those
// steps, so the result is unchanged.  What this buys: one divergent exit per
// K steps instead of one per step, and with it none of the per-step copies of
// the exit state that a divergent exit costs.  An iteration only starts when
// steps + K < limit (limit <= max_steps), so the step cap falls inside the
// replay.  Each iteration also requests the free-map word of its start
// block and, when that region is free for the ray's octant (bitmap at byte
// `fofs` of `fm`), ends the run with kSkipFree: the ray can hit nothing
// more.  Returns kSkipBlocked when it stopped at an iteration that met a
// nonzero byte or left the grid, kSkipLimit at the limit; either way `s` is
// a state the exact loop (or another call) continues from.
template <int K, bool kFree>
__device__ __forceinline__ int dense_skip_empty(float Px, float Py, float Pz, float Ax, float Ay,
                                                float Az, float dx, float dy, float dz, float ix,
                                                float iy, float iz, uint32_t mx, uint32_t my,
                                                uint32_t mz, uint32_t flip, DenseCell& s,
                                                int& steps, int limit,
                                                __amdgpu_buffer_rsrc_t bytes,
                                                __amdgpu_buffer_rsrc_t fm, uint32_t fofs) {
  float bx = s.wx, by = s.wy, bz = s.wz;
  uint32_t bc = s.c;
  int sb = steps;
  int status = kSkipLimit;
  while (sb + K < limit) {
    const uint32_t g = bc ^ flip;
    const uint32_t fw = kFree ? free_word(fm, fofs, g) : 0u;
    DenseCell n;
    n.wx = bx;
    n.wy = by;
    n.wz = bz;
    n.c = bc;
    uint32_t any = 0u;
    bool wrapped = false;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      n = dense_step(Px, Py, Pz, Ax, Ay, Az, dx, dy, dz, ix, iy, iz, mx, my, mz, flip, n);
      any |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(bytes, (int)n.a, 0, 0);
      wrapped |= n.wrapped;
    }
    if (kFree && free_bit(fw, g)) {
      status = kSkipFree;
      break;
    }
    if (wrapped | (any != 0u)) {
      status = kSkipBlocked;
      break;
    }
    bx = n.wx;
    by = n.wy;
    bz = n.wz;
    bc = n.c;
    sb += K;
  }
  s.wx = bx;
  s.wy = by;
  s.wz = bz;
  s.c = bc;
  steps = sb;
  return status;
}

// The empty-space phase for rays whose direction is positive on every axis
// (after the zero substitution) -- the shadow rays of a sun with no negative
// component.  Then rayPositivity is 1 everywhere: the crossing position is
// 8, the re-anchored within is 0 and the Morton code is the address (no
// complemented axes).  Same contract as dense_skip_empty; what differs is how
// the lookup of a state S_i is made exact without a branch:
//   pos >> 3 = G + (int(within) >> 3) moves the probed block one step along
//   each axis k with within_k >= 8, and with positivity 1 that is exactly
//   t_k = (8 - within_k) / d_k <= 0 for the next step.  When one axis has
//   t <= 0 it is the smallest crossing time, so the next step crosses that
//   axis and lands in the very block the corrected lookup probes: the lookup
//   of S_i is the block of S_{i+1}.  Two or more axes at t <= 0 (the median
//   crossing time <= 0) stop the iteration, and the exact loop replays it.
// The lookup of an iteration's last state has no next step inside the
// iteration and takes the general branch (dense_step's).  Shadow rays start
// on lattice points, so this correction is frequent for them: as a branch it
// ran in about half of the wave-steps.
// The free-map test is dense_skip_empty's, on octant 0's bitmap (`fm`).
template <int K, bool kFree>
__device__ __forceinline__ int dense_skip_positive(float dx, float dy, float dz, float ix,
                                                   float iy, float iz, uint32_t mx, uint32_t my,
                                                   uint32_t mz, DenseCell& s, int& steps,
                                                   int limit, __amdgpu_buffer_rsrc_t bytes,
                                                   __amdgpu_buffer_rsrc_t fm) {
  float bx = s.wx, by = s.wy, bz = s.wz;
  uint32_t bc = s.c;
  int sb = steps;
  int status = kSkipLimit;
  while (sb + K < limit) {
    const uint32_t fw = kFree ? free_word(fm, 0u, bc) : 0u;
    float wx = bx, wy = by, wz = bz;
    uint32_t c = bc;
    uint32_t any = 0u;
    bool stop = false;
    float med = 1.0f;  // smallest median crossing time of the iteration
#pragma unroll
    for (int i = 0; i < K; ++i) {
      // the step from S_i (S_0 is the iteration's start, already probed)
      const float tx = (8.0f - wx) * ix;
      const float ty = (8.0f - wy) * iy;
      const float tz = (8.0f - wz) * iz;
      const bool xy = tx < ty, xz = tx < tz, yz = ty < tz;
      const bool m0 = xy & xz;
      const bool m1 = yz & !m0;
      const float tm = fminf(fminf(tx, ty), tz);
      const float nwx = wx + dx * tm;
      const float nwy = wy + dy * tm;
      const float nwz = wz + dz * tm;
      wx = m0 ? 0.0f : nwx;
      wy = m1 ? 0.0f : nwy;
      wz = (m0 | m1) ? nwz : 0.0f;
      const uint32_t M = m0 ? mx : (m1 ? my : mz);
      const uint32_t cm = c & M;
      stop |= cm == M;
      const uint32_t cn = ((cm - M) & M) | (c & ~M);
      if (i > 0) {  // the lookup of S_i
        any |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(bytes, (int)(tm <= 0.0f ? cn : c), 0,
                                                             0);
        med = fminf(med, __builtin_amdgcn_fmed3f(tx, ty, tz));
      }
      c = cn;
    }
    stop |= med <= 0.0f;
    {  // the lookup of S_K
      uint32_t a = c;
      if (fmaxf(fmaxf(wx, wy), wz) >= 8.0f) {
        if (wx >= 8.0f) a = morton_inc_or_out(a, mx);
        if (wy >= 8.0f) a = morton_inc_or_out(a, my);
        if (wz >= 8.0f) a = morton_inc_or_out(a, mz);
      }
      any |= (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(bytes, (int)a, 0, 0);
    }
    if (kFree && free_bit(fw, bc)) {
      status = kSkipFree;
      break;
    }
    if (stop | (any != 0u)) {
      status = kSkipBlocked;
      break;
    }
    bx = wx;
    by = wy;
    bz = wz;
    bc = c;
    sb += K;
  }
  s.wx = bx;
  s.wy = by;
  s.wz = bz;
  s.c = bc;
  steps = sb;
  return status;
}

// The flat lookup's traversal from r's block-level state: the empty-space
// phase, then the exact loop to the ray's end.
// kPositive: the caller guarantees a direction positive on every axis (after
// the zero substitution), and the empty-space phase is dense_skip_positive.
template <class L, bool kPositive = false>
__device__ __forceinline__ void dense_run_prefetch(DenseRay& r, L& lk, const DenseWorld& w,
                                                   DenseHit& h) {
  const uint32_t mx = in_vgpr(w.mx), my = in_vgpr(w.my), mz = in_vgpr(w.mz), flip = r.flip;
  const int max_steps = r.max_steps;
  DenseCell s0;
  s0.wx = r.wx;
  s0.wy = r.wy;
  s0.wz = r.wz;
  s0.c = r.c;
  const __amdgpu_buffer_rsrc_t vox =
      __builtin_amdgcn_make_buffer_rsrc((void*)w.vox, (short)0, (int)w.nvox, 0x00020000);
  int steps0 = r.steps;
  const __amdgpu_buffer_rsrc_t fm = free_rsrc(w.freemap, w.free_bytes);
  int status;
  if constexpr (kPositive)
    status = dense_skip_positive<L::kSkip, L::kFree>(r.dx, r.dy, r.dz, r.ix, r.iy, r.iz, mx,
                                                          my, mz, s0, steps0, max_steps, vox, fm);
  else {
    // the lookups that can only read empty first (primary rays), the
    // iteration that reaches the cap left to the loops below
    const int sblind = steps0;
    if (r.blind > steps0)
      dense_skip_blind<L::kSkip>(r.Px, r.Py, r.Pz, r.Ax, r.Ay, r.Az, r.dx, r.dy, r.dz, r.ix,
                                 r.iy, r.iz, mx, my, mz, flip, s0, steps0,
                                 min(r.blind, max_steps - 1));
    r.blind = steps0 - sblind;  // the steps taken without lookups
    status = dense_skip_empty<L::kSkip, L::kFree>(
        r.Px, r.Py, r.Pz, r.Ax, r.Ay, r.Az, r.dx, r.dy, r.dz, r.ix, r.iy, r.iz, mx, my, mz, flip,
        s0, steps0, max_steps, vox, fm, free_octant_offset(flip, mx, my, mz, w.free_oct_bytes));
  }
  if (L::kFree && status == kSkipFree) {  // a miss: no later lookup can be nonzero
    h.hit = false;
    h.steps = steps0 + L::kSkip;  // the steps executed, the last iteration's included
    return;
  }
  float fwx, fwy, fwz;
  uint32_t fc, fM, fa, fv;
  int sF;
  bool wrappedF;
  dense_loop_prefetch(r.Px, r.Py, r.Pz, r.Ax, r.Ay, r.Az, r.dx, r.dy, r.dz, r.ix, r.iy, r.iz, mx,
                      my, mz, flip, s0, steps0, max_steps, vox, fwx, fwy, fwz, fc, fM, fa, fv,
                      sF, wrappedF);
  (void)lk;
  if (sF >= max_steps || wrappedF) {  // the loop cap, then the bounds test
    h.hit = false;
    h.steps = sF;
    return;
  }
  const uint32_t g = fc ^ flip;
  // faceId = 2 minIdx + 2 - rayPositivity[minIdx] (map.glsl:119-125)
  const bool fx = fM == mx, fy = fM == my;
  const int pm = (flip & fM) ? 0 : 1;  // the complemented axes are the negative ones
  h.hit = true;
  h.face = (fx ? 2 : (fy ? 4 : 6)) - pm;
  h.bx = (int)compact1by2(fa);
  h.by = (int)compact1by2(fa >> 1);
  h.bz = (int)compact1by2(fa >> 2);
  h.code = fa;
  h.v = fv;
  h.sub = (uint32_t)((int)fwx & 7) | ((uint32_t)((int)fwy & 7) << 3) |
          ((uint32_t)((int)fwz & 7) << 6);
  h.hx = (float)(compact1by2(g) << 3) + fwx;
  h.hy = (float)(compact1by2(g >> 1) << 3) + fwy;
  h.hz = (float)(compact1by2(g >> 2) << 3) + fwz;
  h.steps = sF + 1;
}

// kStats with the free-region lookup keeps the fast loops: its statistics
// count the steps executed, rays ending early included.
template <bool kStats, class L>
__device__ __forceinline__ void dense_run(DenseRay& r, L& lk, const DenseWorld& w, DenseHit& h,
                                          uint32_t& l1f, uint32_t& l0f) {
  if constexpr (L::kPrefetch && (!kStats || L::kFree)) {
    dense_run_prefetch(r, lk, w, h);
    return;
  }
  const float Px = r.Px, Py = r.Py, Pz = r.Pz, Ax = r.Ax, Ay = r.Ay, Az = r.Az;
  const float dx = r.dx, dy = r.dy, dz = r.dz, ix = r.ix, iy = r.iy, iz = r.iz;
  const uint32_t mx = w.mx, my = w.my, mz = w.mz, flip = r.flip;
  const int max_steps = r.max_steps;
  float wx = r.wx, wy = r.wy, wz = r.wz;
  uint32_t c = r.c;
  uint32_t M = mx;
  uint32_t a;
  int steps = r.steps;
  uint32_t v;
  bool wrapped;
  do {
    float tx = (Px - wx) * ix;
    float ty = (Py - wy) * iy;
    float tz = (Pz - wz) * iz;
    bool xy = tx < ty, xz = tx < tz, yz = ty < tz;
    bool m0 = xy & xz;
    bool m1 = (!xy) & yz;
    float tm = m0 ? tx : (m1 ? ty : tz);
    float nwx = wx + dx * tm;
    float nwy = wy + dy * tm;
    float nwz = wz + dz * tm;
    wx = m0 ? Ax : nwx;
    wy = m1 ? Ay : nwy;
    wz = (m0 | m1) ? nwz : Az;
    M = m0 ? mx : (m1 ? my : mz);
    const uint32_t cm = c & M;
    wrapped = cm == M;
    c = ((cm - M) & M) | (c & ~M);
    steps++;
    a = c ^ flip;
    bool outside = false;
    if (fmaxf(fmaxf(wx, wy), wz) >= 8.0f) {  // pos >> 3 = G + (int(within) >> 3)
      if (wx >= 8.0f) a = morton_inc(a, mx, outside);
      if (wy >= 8.0f) a = morton_inc(a, my, outside);
      if (wz >= 8.0f) a = morton_inc(a, mz, outside);
    }
    v = outside ? 0u : lk.template probe<kStats>(w, a, l1f, l0f);
  } while (!wrapped & (v == 0u) & (steps < max_steps));

  if (steps >= max_steps || wrapped) {  // the loop cap, then the bounds test
    h.hit = false;
    h.steps = steps;
    return;
  }
  const uint32_t g = c ^ flip;
  h.hit = true;
  h.face = (M == mx ? 2 : (M == my ? 4 : 6)) - ((flip & M) ? 0 : 1);
  h.bx = (int)compact1by2(a);
  h.by = (int)compact1by2(a >> 1);
  h.bz = (int)compact1by2(a >> 2);
  h.code = a;
  h.v = v;
  h.sub = (uint32_t)((int)wx & 7) | ((uint32_t)((int)wy & 7) << 3) |
          ((uint32_t)((int)wz & 7) << 6);
  h.hx = (float)(compact1by2(g) << 3) + wx;
  h.hy = (float)(compact1by2(g >> 1) << 3) + wy;
  h.hz = (float)(compact1by2(g >> 2) << 3) + wz;
  h.steps = steps + 1;
}

// ---------------------------------------------------------------------------
// reference brickmap + atlas layout (literal two-level traceMap)
// ---------------------------------------------------------------------------
struct BrickmapWorld {
  const uint32_t* chunks;  // (dim/8)^3, 0 = empty else pool index + 1
  const uint32_t* pool;    // 512 u32 per block, in-block index x + 8y + 64z
  const uint32_t* atlas;   // 256^3 RGBA8 texels
  uint32_t dim;            // MAP_DIMENSION
  // block occupancy for the empty-space loop (power-of-two dim only, else
  // null): byte at morton(block) = (map_getVoxel(block) != 0)
  const uint8_t* occ;
  uint32_t nvox;           // dim^3
  uint32_t mx, my, mz;     // dilated Morton masks of x, y, z
  // free-region map of the occupancy bytes; free_bytes 0 = none
  const uint32_t* freemap;
  uint32_t free_bytes, free_oct_bytes;
  // the box of the nonempty blocks widened by the traversal's reach (as
  // DenseWorld's: line_meets_reach)
  float reach_lo[3], reach_hi[3];
};

// line_meets_reach of a reference-layout world: the same exactness argument
// (every probe before the first nonempty block is a block-level step of the
// same recurrence); d is the direction before the zero substitution.  Also
// the ray's blind_steps.
__device__ __forceinline__ bool line_meets_reach_bm(const BrickmapWorld& w, float ex, float ey,
                                                    float ez, float dx, float dy, float dz,
                                                    int& blind) {
  const float d[3] = {dx == 0.0f ? 0.001f : dx, dy == 0.0f ? 0.001f : dy,
                      dz == 0.0f ? 0.001f : dz};
  const float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
  blind = blind_steps_box(w.reach_lo, w.reach_hi, ex, ey, ez, d, inv);
  return line_meets_box(w.reach_lo, w.reach_hi, ex, ey, ez, inv);
}

// map_getVoxel (map.glsl:38-47) of block coordinates b; counts the chunk-table
// and pool reads it makes.
template <bool kStats = false>
__device__ __forceinline__ uint32_t bm_block(const BrickmapWorld& w, int bx, int by, int bz,
                                             uint32_t* chunk_reads = nullptr,
                                             uint32_t* pool_reads = nullptr) {
  int M = (int)(w.dim >> 3);
  int cx = bx >> 3, cy = by >> 3, cz = bz >> 3;
  if (cx < 0 || cy < 0 || cz < 0 || cx >= M || cy >= M || cz >= M) return 0;
  uint32_t idx = w.chunks[cx + M * (cy + cz * M)];
  if (kStats) ++*chunk_reads;
  if (idx == 0) return 0;
  size_t o = (size_t)(idx - 1u) * 512u + (size_t)(bx % 8) + (size_t)(((bz % 8) * 8 + (by % 8)) * 8);
  if (kStats) ++*pool_reads;
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
                              float dy, float dz, int max_steps, GeneralHit& h,
                              uint32_t& chunk_reads, uint32_t& pool_atlas_reads) {
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
    uint32_t block = bm_block<kStats>(w, b[0], b[1], b[2], &chunk_reads, &pool_atlas_reads);
    if (block != 0u) {
      uint32_t sub = bm_sub(w, block, p[0] % 8, p[1] % 8, p[2] % 8);
      if (kStats) pool_atlas_reads++;
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


// traceMap over the reference layout with the empty-space steps on the fast
// path: whenever the literal loop is at the block level (stepSize 3) and the
// block it just looked up is empty, the following block-level steps run in
// the pipelined Morton loop over the occupancy bytes (the same recurrence,
// the same stopping rule) until a step lands on a nonempty block, leaves the
// map or reaches the cap; the literal loop then resumes at that block's
// lookup with the same grid / within / minIdx / step count.  Same hits, steps
// and positions as trace_general.  The literal loop is an inner loop that a
// lane leaves when it reaches an empty block: a wave's lanes that leave a
// model block (or start inside one, as shadow rays do) after different
// numbers of sub-voxel steps then run their empty-space steps together
// instead of one group after another.
// kPositive: the caller guarantees a direction with no negative component (the
// zero substitution below makes it positive), and the empty-space steps take
// dense_skip_positive.
// kFree: the empty-space loop tests the free-region map of the occupancy
// bytes and ends a ray that can hit nothing more; the launcher takes it for
// long step caps, so it also takes the longer empty-space iteration.
// blind: leading steps whose lookups can only read empty (blind_steps; 0
// for none).
template <bool kPositive = false, bool kFree = false>
__device__ bool trace_brickmap_fast(const BrickmapWorld& w, float ox, float oy, float oz, float dx,
                                    float dy, float dz, int max_steps, GeneralHit& h,
                                    int blind = 0) {
  constexpr int kSkip = kFree ? kSkipLong : kSkipShort;
  if (dx == 0.0f) dx = 0.001f;
  if (dy == 0.0f) dy = 0.001f;
  if (dz == 0.0f) dz = 0.001f;
  const int B = (int)(8u * w.dim);
  int sgn[3] = {dx > 0.0f ? 1 : (dx < 0.0f ? -1 : 0), dy > 0.0f ? 1 : (dy < 0.0f ? -1 : 0),
                dz > 0.0f ? 1 : (dz < 0.0f ? -1 : 0)};
  int pv[3] = {(1 + sgn[0]) >> 1, (1 + sgn[1]) >> 1, (1 + sgn[2]) >> 1};
  float d[3] = {dx, dy, dz};
  float inv[3] = {1.0f / dx, 1.0f / dy, 1.0f / dz};
  // block-level constants of the fast loop (stepSize 3)
  const float P3[3] = {(float)(pv[0] << 3), (float)(pv[1] << 3), (float)(pv[2] << 3)};
  const float A3[3] = {(float)((1 - pv[0]) << 3) * 0.999f, (float)((1 - pv[1]) << 3) * 0.999f,
                       (float)((1 - pv[2]) << 3) * 0.999f};
  const uint32_t flip = (pv[0] ? 0u : w.mx) | (pv[1] ? 0u : w.my) | (pv[2] ? 0u : w.mz);
  const __amdgpu_buffer_rsrc_t occ =
      __builtin_amdgcn_make_buffer_rsrc((void*)w.occ, (short)0, (int)w.nvox, 0x00020000);
  int m = 0;
  float o8[3] = {ox * 8.0f, oy * 8.0f, oz * 8.0f};
  int g[3] = {f2i_sat(o8[0]), f2i_sat(o8[1]), f2i_sat(o8[2])};
  float wc[3] = {o8[0] - (float)g[0], o8[1] - (float)g[1], o8[2] - (float)g[2]};
  int s = 0;
  h.data = 0;
  h.hx = h.hy = h.hz = -1.0f;
  int step = 0;
  bool hit = false;
  for (;;) {
    // The literal loop (both levels) until the lane stands at an empty
    // block at the block level, hits, leaves the map or reaches the cap.
    // Lanes at an empty block wait here for the rest of the wave, so the
    // wave's empty-space runs below start together.
    bool fast = false;
    for (;;) {
      if (step >= max_steps || g[0] >= B || g[1] >= B || g[2] >= B || g[0] < 0 || g[1] < 0 ||
          g[2] < 0)
        break;
      int p[3] = {g[0] + (int)wc[0], g[1] + (int)wc[1], g[2] + (int)wc[2]};
      int b[3] = {p[0] >> 3, p[1] >> 3, p[2] >> 3};
      uint32_t block = bm_block(w, b[0], b[1], b[2]);
      if (block != 0u) {
        uint32_t sub = bm_sub(w, block, p[0] % 8, p[1] % 8, p[2] % 8);
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
          hit = true;
          break;
        }
        if (s != 0) {
#pragma unroll
          for (int i = 0; i < 3; i++) {
            g[i] += (int)wc[i];
            wc[i] = wc[i] - floorf(wc[i]);
          }
          s = 0;
        }
      } else {
        if (s != 3) {
#pragma unroll
          for (int i = 0; i < 3; i++) {
            int lo = g[i] & 7;
            wc[i] = wc[i] + (float)lo;
            g[i] -= lo;
          }
          s = 3;
        }
        fast = true;  // this iteration's step onwards run in the loop below
        break;
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
      step++;
    }
    if (!fast) break;
    // empty block at the block level: the fast loop takes the next steps
    DenseCell s0;
    s0.wx = wc[0];
    s0.wy = wc[1];
    s0.wz = wc[2];
    s0.c = morton3((uint32_t)g[0] >> 3, (uint32_t)g[1] >> 3, (uint32_t)g[2] >> 3) ^ flip;
    const uint32_t mx = in_vgpr(w.mx), my = in_vgpr(w.my), mz = in_vgpr(w.mz);
    int s0steps = step;
    const int lim = max_steps;
    // the free-region map of the occupancy bytes: block-level lookups only
    // descend into nonempty blocks, so a ray whose block-level lookups all
    // read empty from here on is a miss
    const __amdgpu_buffer_rsrc_t fm = free_rsrc(w.freemap, w.free_bytes);
    int status;
    if constexpr (kPositive)
      status = dense_skip_positive<kSkip, kFree>(d[0], d[1], d[2], inv[0], inv[1], inv[2],
                                                         mx, my, mz, s0, s0steps, lim, occ, fm);
    else {
      if (blind > s0steps)
        dense_skip_blind<kSkip>(P3[0], P3[1], P3[2], A3[0], A3[1], A3[2], d[0], d[1], d[2],
                                inv[0], inv[1], inv[2], mx, my, mz, flip, s0, s0steps,
                                min(blind, max_steps - 1));
      status = dense_skip_empty<kSkip, kFree>(
          P3[0], P3[1], P3[2], A3[0], A3[1], A3[2], d[0], d[1], d[2], inv[0], inv[1], inv[2], mx,
          my, mz, flip, s0, s0steps, lim, occ, fm,
          free_octant_offset(flip, mx, my, mz, w.free_oct_bytes));
    }
    if (kFree && status == kSkipFree) {
      step = s0steps;
      break;
    }
    float fw[3];
    uint32_t fc, fM, fa, fv;
    int sF;
    bool wrappedF;
    dense_loop_prefetch(P3[0], P3[1], P3[2], A3[0], A3[1], A3[2], d[0], d[1], d[2], inv[0],
                        inv[1], inv[2], mx, my, mz, flip, s0, s0steps, max_steps, occ, fw[0],
                        fw[1], fw[2], fc, fM, fa, fv, sF, wrappedF);
    (void)fa;
    (void)fv;
    step = sF;  // iteration sF starts with its bounds test and lookup
    if (wrappedF) break;  // left the map: the bounds test of iteration sF
    const uint32_t gc = fc ^ flip;
    g[0] = (int)(compact1by2(gc) << 3);
    g[1] = (int)(compact1by2(gc >> 1) << 3);
    g[2] = (int)(compact1by2(gc >> 2) << 3);
    wc[0] = fw[0];
    wc[1] = fw[1];
    wc[2] = fw[2];
    m = fM == mx ? 0 : (fM == my ? 1 : 2);
  }
  if (!hit) h.steps = step;
  return hit;
}

}  // namespace vt
