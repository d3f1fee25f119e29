// vt_pixel.hpp -- per-pixel work shared by the render kernels: the primary
// ray, the shadow ray of a hit, the entity test and every store, plus the
// statistics counters and the XCD-aware block remap.
#pragma once
#include "vt_internal.hpp"

namespace vt {

// Block id -> image block.  Blocks are dealt round-robin over the 8 XCDs
// (hardware id b runs on XCD b % 8); each XCD should see runs of neighbouring
// image blocks (their rays share cells in that XCD's L2) without one XCD
// getting a much costlier part of the image than another (a band of sky
// costs a few steps per ray, a band of near terrain hundreds).  So the image
// blocks, in row-major order, are cut into strips of kXcdStrip blocks dealt
// to the XCDs in turn: XCD x traces strips x, x + 8, x + 16, ...  Ids past the
// last full round of 8 strips map to themselves.  Strips of 4 blocks (at
// 8x8 tiles, measured: at 1080p 0.772 ms against 0.775 for 1, 0.773 for 2,
// 0.783 for 8 and 0.792 for 16; the engine's own frame and the reference
// layout gain 2-3% the same way; clocks settled, one box).
constexpr uint32_t kXcdStrip = 4u;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t n) {
  const uint32_t X = 8u, round = X * kXcdStrip;
  if (bid >= (n / round) * round) return bid;
  const uint32_t x = bid % X, j = bid / X;
  return ((j / kXcdStrip) * X + x) * kXcdStrip + j % kXcdStrip;
}

struct Primary {
  bool hit;
  int face;
  uint32_t voxel;  // linear block id
  uint32_t data;   // RGBA8 colour (packUnorm4x8 texel)
  uint32_t sub;
  uint32_t material;
  uint32_t code;   // Morton code of the hit block (dense worlds)
  float hx, hy, hz;
  int steps;
};

struct Counters {
  uint32_t psteps, ssteps, hits, occluded, l1f, l0f, leaf, ent;
};

// ---- Lambert x visibility shading and its gradients (project-defined) ----
struct LambertSun {
  float lx, ly, lz, ls;  // sun / |sun| and |sun|
};

__device__ __forceinline__ LambertSun lambert_sun(const float sun[3]) {
  LambertSun s;
  s.ls = len3(sun[0], sun[1], sun[2]);
  s.lx = sun[0] / s.ls;
  s.ly = sun[1] / s.ls;
  s.lz = sun[2] / s.ls;
  return s;
}

__device__ __forceinline__ float lambert_nd(int face, const LambertSun& s) {
  float nx, ny, nz;
  face_normal(face, nx, ny, nz);
  float nd = nx * s.lx;
  nd = nd + ny * s.ly;
  nd = nd + nz * s.lz;
  return nd;
}

__device__ __forceinline__ float lambert_factor(float ambient, float nd, bool occluded) {
  return ambient + fmaxf(nd, 0.0f) * (occluded ? 0.0f : 1.0f);
}

// Compact albedo index of the voxel with Morton code `code`: occupied voxels
// before its 4^3 brick plus those before it inside the brick.
__device__ __forceinline__ uint32_t compact_index(const DenseWorld& w, uint32_t code) {
  const uint32_t word = code >> 6, bit = code & 63u;
  const uint64_t below = bit ? (w.l0[word] & (~0ull >> (64u - bit))) : 0ull;
  return w.prefix[word] + (uint32_t)__popcll(below);
}

__device__ __forceinline__ uint32_t voxel_code(const DenseWorld& w, uint32_t voxel) {
  const uint32_t m = w.dim - 1u;
  return morton3(voxel & m, (voxel >> w.log2dim) & m, voxel >> (2u * w.log2dim));
}

__device__ __forceinline__ uint32_t albedo_index(const DenseWorld& w, uint32_t flags,
                                                 uint32_t voxel, uint32_t code) {
  return (flags & VT_FLAG_ALBEDO_COMPACT) ? compact_index(w, code) : voxel;
}

struct GradAcc {
  float loss, sx, sy, sz;  // per-lane partial sums, flushed per wave
  uint32_t key;            // albedo index of the pending albedo gradient
  float ga[3];             // pending dL/dalbedo of this lane's pixel
};

constexpr uint32_t kNoKey = 0xFFFFFFFFu;

__device__ __forceinline__ GradAcc grad_acc() {
  GradAcc a;
  a.loss = a.sx = a.sy = a.sz = 0.0f;
  a.key = kNoKey;
  a.ga[0] = a.ga[1] = a.ga[2] = 0.0f;
  return a;
}

// Gradients of one hit pixel with dL/dI = g: dL/dA_c = g_c * fct (left
// pending in acc until flush_albedo), and, when lit, dL/dsun =
// (sum_c g_c A_c) * (n - (n.l) l) / |sun|.
__device__ __forceinline__ void lambert_accumulate(uint32_t aidx, int face, float nd,
                                                   bool occluded, const float* A, const float g[3],
                                                   float fct, const LambertSun& s, GradAcc& acc) {
  acc.key = aidx;
  acc.ga[0] = g[0] * fct;
  acc.ga[1] = g[1] * fct;
  acc.ga[2] = g[2] * fct;
  if (nd > 0.0f && !occluded) {
    const float k = (g[0] * A[0] + g[1] * A[1] + g[2] * A[2]) / s.ls;
    float nx, ny, nz;
    face_normal(face, nx, ny, nz);
    acc.sx += k * (nx - nd * s.lx);
    acc.sy += k * (ny - nd * s.ly);
    acc.sz += k * (nz - nd * s.lz);
  }
}

__device__ __forceinline__ float wave_sum_f(float x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Segmented sum over runs of equal keys in consecutive lanes (neighbouring
// pixels of a row usually hit one voxel): on return the first lane of each
// run holds the run's sums and gets true.  Every lane of the wave must call.
__device__ __forceinline__ bool wave_run_sum(uint32_t key, float& a, float& b, float& c) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0u || prev != key;
  const uint64_t heads = __ballot(head);
  const uint32_t run = (uint32_t)__popcll(heads & (~0ull >> (63u - lane)));
#pragma unroll
  for (uint32_t off = 1; off < 64u; off <<= 1) {
    const float a2 = __shfl_down(a, off, 64), b2 = __shfl_down(b, off, 64),
                c2 = __shfl_down(c, off, 64);
    const uint32_t r2 = __shfl_down(run, off, 64);
    if (lane + off < 64u && r2 == run) {
      a += a2;
      b += b2;
      c += c2;
    }
  }
  return head;
}

// Adds the wave's pending albedo gradients: one atomic triple per run of
// lanes with the same albedo index.  Every lane of the wave must call this.
__device__ __forceinline__ void flush_albedo(const ShadeConsts& sc, GradAcc& acc) {
  float a = acc.ga[0], b = acc.ga[1], c = acc.ga[2];
  const bool head = wave_run_sum(acc.key, a, b, c);
  if (head && acc.key != kNoKey) {
    float* g = sc.grad_albedo + 3 * (size_t)acc.key;
    atomicAdd(g + 0, a);
    atomicAdd(g + 1, b);
    atomicAdd(g + 2, c);
  }
  acc.key = kNoKey;
  acc.ga[0] = acc.ga[1] = acc.ga[2] = 0.0f;
}

// Every lane of the wave must call this.
// Loss and dL/dsun of the whole block: wave sums, then one sum over the
// block's waves in LDS, then one atomic per quantity and block.  Float
// atomics on one address serialise in the memory system, so they are kept
// to a few thousand per launch.  Every thread of the block must call this.
__device__ __forceinline__ void flush_grad(const ShadeConsts& sc, const GradAcc& acc) {
  __shared__ float part[4][4];  // [quantity][wave], 256-thread blocks
  const float v[4] = {wave_sum_f(acc.loss), wave_sum_f(acc.sx), wave_sum_f(acc.sy),
                      wave_sum_f(acc.sz)};
  const uint32_t wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0u) {
#pragma unroll
    for (int q = 0; q < 4; q++) part[q][wave] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 4u) {
    const uint32_t q = threadIdx.x;
    const float t = (part[q][0] + part[q][1]) + (part[q][2] + part[q][3]);
    float* dst = q == 0u ? sc.loss : (sc.grad_sun ? sc.grad_sun + (q - 1u) : nullptr);
    if (dst) atomicAdd(dst, t);
  }
}

// Shadow origin: position + normal * 0.001 with the values the G-buffer
// round trip would give (secondary.comp.glsl:25,36-37).
__device__ __forceinline__ void shadow_origin(const Primary& p, uint32_t flags, float& sox,
                                              float& soy, float& soz) {
  float px = ceilf(p.hx) / 8.0f, py = ceilf(p.hy) / 8.0f, pz = ceilf(p.hz) / 8.0f;
  float nx, ny, nz;
  face_normal(p.face, nx, ny, nz);
  if (flags & VT_FLAG_GBUF_UNORM8) {
    nx = unorm8(nx);
    ny = unorm8(ny);
    nz = unorm8(nz);
  }
  sox = px + nx * 0.001f;
  soy = py + ny * 0.001f;
  soz = pz + nz * 0.001f;
}

// `dw` is the dense world (Lambert shading reads its compact index), null for
// the reference layout (which the host never runs with VT_FLAG_LAMBERT).
// Shading model of a launch, fixed at compile time so the reference composite
// carries no Lambert or gradient code (and no registers for it).
enum ShadeMode : int { kShadeRef = 0, kShadeLambert = 1, kShadeLambertGrad = 2 };

__host__ __device__ constexpr int shade_mode(uint32_t flags) {
  return (flags & VT_FLAG_GRAD_L2) ? kShadeLambertGrad
                                   : ((flags & VT_FLAG_LAMBERT) ? kShadeLambert : kShadeRef);
}

// Colour image of a hit: unpackUnorm4x8(data) (primary.comp.glsl:58-60).
__device__ __forceinline__ float4 hit_color(uint32_t data, bool unorm) {
  float4 col = make_float4((float)(data & 255u) / 255.0f, (float)((data >> 8) & 255u) / 255.0f,
                           (float)((data >> 16) & 255u) / 255.0f, (float)(data >> 24) / 255.0f);
  if (unorm) col = make_float4(unorm8(col.x), unorm8(col.y), unorm8(col.z), unorm8(col.w));
  return col;
}

// Normal and position images and the hit distance of a primary hit
// (primary.comp.glsl:61-63).
__device__ __forceinline__ void hit_normal_position(const ViewConsts& v, const Primary& p,
                                                    bool unorm, float4& nrm, float4& pos,
                                                    float& t) {
  float nx, ny, nz;
  face_normal(p.face, nx, ny, nz);
  nrm = make_float4(nx, ny, nz, 1.0f);
  if (unorm) nrm = make_float4(unorm8(nrm.x), unorm8(nrm.y), unorm8(nrm.z), unorm8(nrm.w));
  pos = make_float4(ceilf(p.hx) / 8.0f, ceilf(p.hy) / 8.0f, ceilf(p.hz) / 8.0f, 1.0f);
  t = len3(p.hx / 8.0f - v.pos[0], p.hy / 8.0f - v.pos[1], p.hz / 8.0f - v.pos[2]);
}

// traceEntities(rayOrigin, rayDir, distance(rayOrigin, inter.hit_pos / 8.))
// and the illumination image of a hit (secondary.comp.glsl:36-50).
__device__ __forceinline__ float4 shadow_illum(const ShadeConsts& sc, bool unorm, bool shit,
                                               float shx, float shy, float shz, float sox,
                                               float soy, float soz, bool& ent, bool& occl) {
  float hx8 = shx / 8.0f, hy8 = shy / 8.0f, hz8 = shz / 8.0f;
  float maxd = len3(sox - hx8, soy - hy8, soz - hz8);
  ent = trace_entities(sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2], maxd);
  occl = ent || shit;
  float4 ill = make_float4(sc.sun[0], sc.sun[1], sc.sun[2], occl ? -0.3f : 0.3f);
  if (unorm) ill = make_float4(unorm8(ill.x), unorm8(ill.y), unorm8(ill.z), unorm8(ill.w));
  return ill;
}

// Hit-record info word: face | shadow << 3 | sub-voxel << 4 | material << 16.
__device__ __forceinline__ uint32_t hit_info(int face, uint32_t sub, uint32_t material,
                                             bool occl) {
  return (uint32_t)face | ((occl ? 1u : 0u) << 3) | (sub << 4) | (material << 16);
}

// Reference composite of a hit (blit.fragment.glsl:30-31), the optional
// present pass, and its stores.
__device__ __forceinline__ void store_ref_composite(const ViewConsts& v, const ShadeConsts& sc,
                                                    const FrameOut& out, size_t pix, uint32_t px,
                                                    uint32_t py, float4 col, float4 ill,
                                                    bool hit) {
  float4 s;
  if (hit) {
    float r, g, b;
    sky(sc.sun, ill.x, ill.y, ill.z, r, g, b);
    s = make_float4(col.x + ill.w * r, col.y + ill.w * g, col.z + ill.w * b,
                    col.w + ill.w * 1.0f);
  } else {
    s = col;
  }
  if (sc.flags & VT_FLAG_PRESENT) {
    s = present_pixel(s, px, py, v.W, v.H);
    if (out.display) out.display[pix] = pack_unorm8(s);
  }
  out.shaded[pix] = s;
}

template <int kShade>
__device__ __forceinline__ void write_pixel(const ViewConsts& v, const ShadeConsts& sc,
                                            const FrameOut& out, uint32_t px, uint32_t py,
                                            const Primary& p, bool shit, float shx, float shy,
                                            float shz, float sox, float soy, float soz,
                                            int ssteps, Counters& cnt, const DenseWorld* dw,
                                            GradAcc& gacc) {
  const size_t pix = (size_t)py * v.W + px;
  const bool unorm = (sc.flags & VT_FLAG_GBUF_UNORM8) != 0u;
  float4 col, nrm, pos, ill;
  float t;
  bool occl = false;
  if (p.hit) {
    col = hit_color(p.data, unorm);
    hit_normal_position(v, p, unorm, nrm, pos, t);
    bool ent;
    ill = shadow_illum(sc, unorm, shit, shx, shy, shz, sox, soy, soz, ent, occl);
    cnt.hits++;
    cnt.occluded += occl ? 1u : 0u;
    cnt.ent += ent ? 1u : 0u;
    cnt.ssteps += (uint32_t)ssteps;
  } else {
    float dx, dy, dz;
    raygen(v, px, py, dx, dy, dz);
    float r, g, b;
    sky(sc.sun, dx, dy, dz, r, g, b);
    col = make_float4(r, g, b, 1.0f);
    if (unorm) col = make_float4(unorm8(r), unorm8(g), unorm8(b), 1.0f);
    nrm = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    pos = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
    ill = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    t = __builtin_inff();
  }
  cnt.psteps += (uint32_t)p.steps;

  if (out.color) out.color[pix] = col;
  if (out.normal) out.normal[pix] = nrm;
  if (out.position) out.position[pix] = pos;
  if (out.illum) out.illum[pix] = ill;
  if (out.steps) out.steps[pix] = (uint32_t)p.steps | ((uint32_t)(p.hit ? ssteps : 0) << 16);
  if (out.hits) {
    uint32_t info = p.hit ? hit_info(p.face, p.sub, p.material, occl) : 0u;
    out.hits[3 * pix + 0] = p.hit ? p.voxel : 0xFFFFFFFFu;
    out.hits[3 * pix + 1] = __float_as_uint(t);
    out.hits[3 * pix + 2] = info;
  }
  if constexpr (kShade != kShadeRef) {
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    constexpr bool grad = kShade == kShadeLambertGrad;
    float4 T = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if constexpr (grad) T = out.target[pix];
    if (p.hit) {
      const LambertSun ls = lambert_sun(sc.sun);
      const float nd = lambert_nd(p.face, ls);
      const float fct = lambert_factor(sc.ambient, nd, occl);
      const uint32_t ai = albedo_index(*dw, sc.flags, p.voxel, p.code);
      const float* a = sc.albedo + 3 * (size_t)ai;
      s = make_float4(a[0] * fct, a[1] * fct, a[2] * fct, 1.0f);
      if constexpr (grad) {
        const float r[3] = {s.x - T.x, s.y - T.y, s.z - T.z};
        gacc.loss += r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        const float g[3] = {2.0f * r[0], 2.0f * r[1], 2.0f * r[2]};
        lambert_accumulate(ai, p.face, nd, occl, a, g, fct, ls, gacc);
      }
    } else if constexpr (grad) {
      gacc.loss += T.x * T.x + T.y * T.y + T.z * T.z;
    }
    if (out.shaded) out.shaded[pix] = s;
  } else if (out.shaded) {
    store_ref_composite(v, sc, out, pix, px, py, col, ill, p.hit);
  }
}

// The reference composite of a primary hit in two halves, so that only what
// the second half needs stays live across the shadow traversal (the rest of
// the primary hit would otherwise be spilled around the shadow loops).
// First half: the stores that do not depend on the shadow ray.
__device__ __forceinline__ void store_hit_primary(const ViewConsts& v, const ShadeConsts& sc,
                                                  const FrameOut& out, size_t pix,
                                                  const Primary& p) {
  const bool unorm = (sc.flags & VT_FLAG_GBUF_UNORM8) != 0u;
  float4 nrm, pos;
  float t;
  hit_normal_position(v, p, unorm, nrm, pos, t);
  if (out.color) out.color[pix] = hit_color(p.data, unorm);
  if (out.normal) out.normal[pix] = nrm;
  if (out.position) out.position[pix] = pos;
  if (out.hits) {
    out.hits[3 * pix + 0] = p.voxel;
    out.hits[3 * pix + 1] = __float_as_uint(t);
  }
}

// Second half, once the shadow ray (shit, sh*, ssteps) is done: data = the
// primary colour, info0 = its hit-record info word without the shadow bit.
__device__ __forceinline__ void store_hit_shadow_ref(const ViewConsts& v, const ShadeConsts& sc,
                                                     const FrameOut& out, uint32_t px,
                                                     uint32_t py, uint32_t data, uint32_t info0,
                                                     int psteps, bool shit, float shx,
                                                     float shy, float shz, float sox, float soy,
                                                     float soz, int ssteps, Counters& cnt) {
  const size_t pix = (size_t)py * v.W + px;
  const bool unorm = (sc.flags & VT_FLAG_GBUF_UNORM8) != 0u;
  bool ent, occl;
  const float4 ill = shadow_illum(sc, unorm, shit, shx, shy, shz, sox, soy, soz, ent, occl);
  cnt.hits++;
  cnt.occluded += occl ? 1u : 0u;
  cnt.ent += ent ? 1u : 0u;
  cnt.ssteps += (uint32_t)ssteps;
  cnt.psteps += (uint32_t)psteps;
  if (out.illum) out.illum[pix] = ill;
  if (out.steps) out.steps[pix] = (uint32_t)psteps | ((uint32_t)ssteps << 16);
  if (out.hits) out.hits[3 * pix + 2] = info0 | ((occl ? 1u : 0u) << 3);
  if (out.shaded)
    store_ref_composite(v, sc, out, pix, px, py, hit_color(data, unorm), ill, true);
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

__device__ __forceinline__ void flush_counters(const FrameOut& out, const Counters& c) {
  uint32_t v[8] = {c.psteps, c.ssteps, c.hits, c.occluded, c.l1f, c.l0f, c.leaf, c.ent};
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint32_t s = wave_sum(v[i]);
    if ((threadIdx.x & 63u) == 0u && s) atomicAdd(out.stats + i, (unsigned long long)s);
  }
}

// One pixel of a dense world: primary ray (primary.comp.glsl:31-68), then the
// shadow ray of a hit (secondary.comp.glsl:25-50), then every store.
template <class L, bool kStats, int kShade>
__device__ __forceinline__ void trace_pixel_dense(const DenseWorld& w, const ViewConsts& v,
                                                  const ShadeConsts& sc, const FrameOut& out,
                                                  uint32_t px, uint32_t py, Counters& cnt,
                                                  GradAcc& gacc) {
  DenseRay r;
  L lk;
  DenseHit h;
  Primary p;
  // ---- primary ray (primary.comp.glsl:31-43) ----
  float dx, dy, dz;
  raygen(v, px, py, dx, dy, dz);
  float tn, tf;
  const float D = (float)w.dim;
  intersect_aabb(v.pos[0], v.pos[1], v.pos[2], dx, dy, dz, 0.0f, 0.0f, 0.0f, D, D, D, tn, tf);
  float tc = fmaxf(tn, 0.0f);
  float ex = (v.pos[0] + dx * tc) - 0.001f;
  float ey = (v.pos[1] + dy * tc) - 0.001f;
  float ez = (v.pos[2] + dz * tc) - 0.001f;
  if (!dense_begin<kStats>(r, lk, w, ex, ey, ez, dx, dy, dz, sc.max_primary, h, cnt.l1f,
                           cnt.l0f))
    dense_run<kStats>(r, lk, w, h, cnt.l1f, cnt.l0f);
  p.hit = h.hit;
  p.steps = h.steps;
  float sox = 0.f, soy = 0.f, soz = 0.f;
  bool shit = false;
  float shx = -1.0f, shy = -1.0f, shz = -1.0f;
  int ssteps = 0;
  if (p.hit) {
    p.face = h.face;
    p.voxel = (uint32_t)h.bx + w.dim * ((uint32_t)h.by + w.dim * (uint32_t)h.bz);
    p.material = lk.material(w, h.code, h.v, cnt.leaf);
    p.data = w.palette[p.material];
    p.sub = h.sub;
    p.code = h.code;
    p.hx = h.hx;
    p.hy = h.hy;
    p.hz = h.hz;
    // ---- shadow ray (secondary.comp.glsl:25-42) ----
    shadow_origin(p, sc.flags, sox, soy, soz);
    if constexpr (kShade == kShadeRef) store_hit_primary(v, sc, out, (size_t)py * v.W + px, p);
    const uint32_t data = p.data, info0 = hit_info(p.face, p.sub, p.material, false);
    const int psteps = p.steps;
    if (!dense_begin_k<kStats>(r, lk, w, sox, soy, soz, sc.shadow, sc.max_shadow, h, cnt.l1f,
                               cnt.l0f)) {
      if constexpr (L::kPrefetch && (!kStats || L::kFree)) {
        // the sun is the same for every lane: a uniform branch
        if (sc.sun[0] >= 0.0f && sc.sun[1] >= 0.0f && sc.sun[2] >= 0.0f)
          dense_run_prefetch<L, true>(r, lk, w, h);
        else
          dense_run_prefetch<L, false>(r, lk, w, h);
      } else {
        dense_run<kStats>(r, lk, w, h, cnt.l1f, cnt.l0f);
      }
    }
    shit = h.hit;
    if (shit) {
      shx = h.hx;
      shy = h.hy;
      shz = h.hz;
    }
    ssteps = h.steps;
    if constexpr (kShade == kShadeRef) {
      store_hit_shadow_ref(v, sc, out, px, py, data, info0, psteps, shit, shx, shy, shz, sox, soy,
                           soz, ssteps, cnt);
      return;
    }
  }
  write_pixel<kShade>(v, sc, out, px, py, p, shit, shx, shy, shz, sox, soy, soz, ssteps, cnt, &w,
                     gacc);
}


}  // namespace vt
