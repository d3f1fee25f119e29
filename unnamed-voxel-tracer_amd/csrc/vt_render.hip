// vt_render.hip -- fused primary + shadow render kernels for gfx950.
//
// One launch does what the reference splits into two dispatches plus a blit:
// raygen + traceMap + G-buffer store (primary.comp.glsl:23-69), shadow
// traceMap + traceEntities (secondary.comp.glsl:18-50) and the composite
// formula (blit.fragment.glsl:30-31).  The G-buffer round trip between the
// two passes stays in registers; the G-buffer images are only written when
// the caller asks for them.
//
// Launch shape: 256-thread blocks = 4 wave64s; each wave owns one tile of 64
// pixels (8x8, 16x4 or 32x2) for ray coherence, a block 2x2 tiles.  Block ids
// are remapped so that each XCD (blocks are dealt round-robin over the 8 XCDs)
// receives a contiguous band of the image, keeping each XCD's L2 working set
// local.
//
// Dense worlds run primary and shadow rays through ONE loop per lane: a lane
// whose primary ray terminates starts its shadow ray in the next iteration
// instead of idling until the whole wave's primary rays are done.
#include "vt_internal.hpp"

namespace vt {

__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t n) {
  const uint32_t X = 8u;
  uint32_t q = n / X, r = n % X;
  uint32_t x = bid % X, j = bid / X;
  return x < r ? x * (q + 1u) + j : r * (q + 1u) + (x - r) * q + j;
}

struct Primary {
  bool hit;
  int face;
  uint32_t voxel;  // linear block id
  uint32_t data;   // RGBA8 colour (packUnorm4x8 texel)
  uint32_t sub;
  uint32_t material;
  float hx, hy, hz;
  int steps;
};

struct Counters {
  uint32_t psteps, ssteps, hits, occluded, l1f, l0f, leaf, ent;
};

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

__device__ __forceinline__ void write_pixel(const ViewConsts& v, const ShadeConsts& sc,
                                            const FrameOut& out, uint32_t px, uint32_t py,
                                            const Primary& p, bool shit, float shx, float shy,
                                            float shz, float sox, float soy, float soz,
                                            int ssteps, Counters& cnt) {
  const size_t pix = (size_t)py * v.W + px;
  const bool unorm = (sc.flags & VT_FLAG_GBUF_UNORM8) != 0u;
  float4 col, nrm, pos, ill;
  float t;
  bool occl = false;
  if (p.hit) {
    col = make_float4((float)(p.data & 255u) / 255.0f, (float)((p.data >> 8) & 255u) / 255.0f,
                      (float)((p.data >> 16) & 255u) / 255.0f, (float)(p.data >> 24) / 255.0f);
    float nx, ny, nz;
    face_normal(p.face, nx, ny, nz);
    nrm = make_float4(nx, ny, nz, 1.0f);
    pos = make_float4(ceilf(p.hx) / 8.0f, ceilf(p.hy) / 8.0f, ceilf(p.hz) / 8.0f, 1.0f);
    if (unorm) {
      col = make_float4(unorm8(col.x), unorm8(col.y), unorm8(col.z), unorm8(col.w));
      nrm = make_float4(unorm8(nrm.x), unorm8(nrm.y), unorm8(nrm.z), unorm8(nrm.w));
    }
    // traceEntities(rayOrigin, rayDir, distance(rayOrigin, inter.hit_pos / 8.))
    float hx8 = shx / 8.0f, hy8 = shy / 8.0f, hz8 = shz / 8.0f;
    float maxd = len3(sox - hx8, soy - hy8, soz - hz8);
    bool ent = trace_entities(sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2], maxd);
    occl = ent || shit;
    ill = make_float4(sc.sun[0], sc.sun[1], sc.sun[2], occl ? -0.3f : 0.3f);
    if (unorm) ill = make_float4(unorm8(ill.x), unorm8(ill.y), unorm8(ill.z), unorm8(ill.w));
    t = len3(p.hx / 8.0f - v.pos[0], p.hy / 8.0f - v.pos[1], p.hz / 8.0f - v.pos[2]);
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
    uint32_t info = p.hit ? ((uint32_t)p.face | ((occl ? 1u : 0u) << 3) | (p.sub << 4) |
                             (p.material << 16))
                          : 0u;
    out.hits[3 * pix + 0] = p.hit ? p.voxel : 0xFFFFFFFFu;
    out.hits[3 * pix + 1] = __float_as_uint(t);
    out.hits[3 * pix + 2] = info;
  }
  if (out.shaded) {
    float4 s;
    if (sc.flags & VT_FLAG_LAMBERT) {
      if (p.hit) {
        float ls = len3(sc.sun[0], sc.sun[1], sc.sun[2]);
        float lx = sc.sun[0] / ls, ly = sc.sun[1] / ls, lz = sc.sun[2] / ls;
        float nx, ny, nz;
        face_normal(p.face, nx, ny, nz);
        float nd = nx * lx;
        nd = nd + ny * ly;
        nd = nd + nz * lz;
        float fct = sc.ambient + fmaxf(nd, 0.0f) * (occl ? 0.0f : 1.0f);
        const float* a = sc.albedo + 3 * (size_t)p.voxel;
        s = make_float4(a[0] * fct, a[1] * fct, a[2] * fct, 1.0f);
      } else {
        s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    } else if (p.hit) {
      float r, g, b;
      sky(sc.sun, ill.x, ill.y, ill.z, r, g, b);
      s = make_float4(col.x + ill.w * r, col.y + ill.w * g, col.z + ill.w * b,
                      col.w + ill.w * 1.0f);
    } else {
      s = col;
    }
    out.shaded[pix] = s;
  }
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

template <class L, bool kStats>
__global__ __launch_bounds__(256) void k_render_dense(DenseWorld w, ViewConsts v, ShadeConsts sc,
                                                      FrameOut out, uint32_t log2_tw,
                                                      uint32_t blocks_x) {
  const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t bx = bid % blocks_x, by = bid / blocks_x;
  const uint32_t px = bx * (2u * tw) + (wave & 1u) * tw + (lane & (tw - 1u));
  const uint32_t py = by * (2u * th) + (wave >> 1) * th + (lane >> log2_tw);
  const bool active = px < v.W && py < v.H;

  Counters cnt = {0, 0, 0, 0, 0, 0, 0, 0};
  if (active) {
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
      p.hx = h.hx;
      p.hy = h.hy;
      p.hz = h.hz;
      // ---- shadow ray (secondary.comp.glsl:25-42) ----
      shadow_origin(p, sc.flags, sox, soy, soz);
      if (!dense_begin<kStats>(r, lk, w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
                               sc.max_shadow, h, cnt.l1f, cnt.l0f))
        dense_run<kStats>(r, lk, w, h, cnt.l1f, cnt.l0f);
      shit = h.hit;
      if (shit) {
        shx = h.hx;
        shy = h.hy;
        shz = h.hz;
      }
      ssteps = h.steps;
    }
    write_pixel(v, sc, out, px, py, p, shit, shx, shy, shz, sox, soy, soz, ssteps, cnt);
  }
  if (kStats) flush_counters(out, cnt);
}

template <bool kStats>
__global__ __launch_bounds__(256) void k_render_brickmap(BrickmapWorld w, ViewConsts v,
                                                         ShadeConsts sc, FrameOut out,
                                                         uint32_t log2_tw, uint32_t blocks_x) {
  const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t bx = bid % blocks_x, by = bid / blocks_x;
  const uint32_t px = bx * (2u * tw) + (wave & 1u) * tw + (lane & (tw - 1u));
  const uint32_t py = by * (2u * th) + (wave >> 1) * th + (lane >> log2_tw);
  const bool active = px < v.W && py < v.H;
  Counters cnt = {0, 0, 0, 0, 0, 0, 0, 0};
  if (active) {
    float dx, dy, dz;
    raygen(v, px, py, dx, dy, dz);
    float tn, tf;
    const float D = (float)w.dim;
    intersect_aabb(v.pos[0], v.pos[1], v.pos[2], dx, dy, dz, 0.0f, 0.0f, 0.0f, D, D, D, tn, tf);
    float tc = fmaxf(tn, 0.0f);
    float ex = (v.pos[0] + dx * tc) - 0.001f;
    float ey = (v.pos[1] + dy * tc) - 0.001f;
    float ez = (v.pos[2] + dz * tc) - 0.001f;
    GeneralHit g;
    trace_general<kStats>(w, ex, ey, ez, dx, dy, dz, sc.max_primary, g, cnt.leaf);
    Primary p;
    p.hit = g.data != 0u;
    p.steps = g.steps;
    float sox = 0.f, soy = 0.f, soz = 0.f;
    bool shit = false;
    float shx = -1.0f, shy = -1.0f, shz = -1.0f;
    int ssteps = 0;
    if (p.hit) {
      p.face = g.face;
      p.voxel = (uint32_t)g.bx + w.dim * ((uint32_t)g.by + w.dim * (uint32_t)g.bz);
      p.data = g.data;
      p.material = bm_block(w, g.bx, g.by, g.bz) & 255u;
      p.sub = g.sub;
      p.hx = g.hx;
      p.hy = g.hy;
      p.hz = g.hz;
      shadow_origin(p, sc.flags, sox, soy, soz);
      GeneralHit s;
      trace_general<kStats>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2], sc.max_shadow, s,
                            cnt.leaf);
      shit = s.data != 0u;
      if (shit) {
        shx = s.hx;
        shy = s.hy;
        shz = s.hz;
      }
      ssteps = s.steps;
    }
    write_pixel(v, sc, out, px, py, p, shit, shx, shy, shz, sox, soy, soz, ssteps, cnt);
  }
  if (kStats) flush_counters(out, cnt);
}

static void grid_shape(const ViewConsts& v, TileShape tile, uint32_t& blocks_x, uint32_t& nblocks) {
  uint32_t tw = 1u << tile.log2_w, th = 64u >> tile.log2_w;
  blocks_x = (v.W + 2u * tw - 1u) / (2u * tw);
  uint32_t blocks_y = (v.H + 2u * th - 1u) / (2u * th);
  nblocks = blocks_x * blocks_y;
}

hipError_t launch_render_dense(const DenseWorld& w, uint32_t traversal, const ViewConsts& v,
                               const ShadeConsts& sc, const FrameOut& out, TileShape tile,
                               hipStream_t stream) {
  uint32_t blocks_x, nblocks;
  grid_shape(v, tile, blocks_x, nblocks);
  dim3 grid(nblocks), block(256);
  const bool stats = out.stats != nullptr;
  if (traversal == VT_TRAVERSAL_BITMASK) {
    if (stats)
      hipLaunchKernelGGL((k_render_dense<BitmaskLookup, true>), grid, block, 0, stream, w, v, sc,
                         out, tile.log2_w, blocks_x);
    else
      hipLaunchKernelGGL((k_render_dense<BitmaskLookup, false>), grid, block, 0, stream, w, v, sc,
                         out, tile.log2_w, blocks_x);
  } else {
    if (stats)
      hipLaunchKernelGGL((k_render_dense<FlatLookup, true>), grid, block, 0, stream, w, v, sc, out,
                         tile.log2_w, blocks_x);
    else
      hipLaunchKernelGGL((k_render_dense<FlatLookup, false>), grid, block, 0, stream, w, v, sc,
                         out, tile.log2_w, blocks_x);
  }
  return hipGetLastError();
}

hipError_t launch_render_brickmap(const BrickmapWorld& w, const ViewConsts& v,
                                  const ShadeConsts& sc, const FrameOut& out, TileShape tile,
                                  hipStream_t stream) {
  uint32_t blocks_x, nblocks;
  grid_shape(v, tile, blocks_x, nblocks);
  dim3 grid(nblocks), block(256);
  if (out.stats)
    hipLaunchKernelGGL((k_render_brickmap<true>), grid, block, 0, stream, w, v, sc, out,
                       tile.log2_w, blocks_x);
  else
    hipLaunchKernelGGL((k_render_brickmap<false>), grid, block, 0, stream, w, v, sc, out,
                       tile.log2_w, blocks_x);
  return hipGetLastError();
}

}  // namespace vt
