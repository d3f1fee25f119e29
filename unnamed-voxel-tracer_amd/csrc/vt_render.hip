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
// are remapped (xcd_remap) so that each XCD receives strips of neighbouring
// blocks spread over the whole image: its L2 working set stays local and
// no XCD is left with the costliest rows.  Each lane traces its primary ray,
// then (on a hit) its shadow ray.
//
// Rays end early where nothing is left to hit: with the free-region test
// (a per-world bitmap of the 16^3 regions from which every region ahead in a
// direction octant is empty) a ray standing in such a region is a miss.  Lit
// shadow rays, which otherwise run to the edge of the grid, mostly end a few
// regions after their start, so the lanes of a wave finish close together.
#include "vt_pixel.hpp"

namespace vt {

// amdgpu_waves_per_eu: at most 72 VGPRs (seven waves per SIMD) for every
// shading mode.  The reference composite needs no scratch at 72 (56 B per
// lane of spills at 64 with the free-region test, ~100 MB of spill stores
// per 1080p frame); C1 at 1080p measures the same at eight, seven and six
// waves (0.2785 / 0.2791 / 0.2789 ms), C3's 512x512 views 2% faster at
// seven (0.1113 vs 0.1137 ms per view).  Before the free-region test eight
// waves were faster (1080p 0.785 vs 0.804 ms): the traversal loop then held
// most of the time and its gathers needed the waves.
#ifndef VT_REF_WAVES_PER_SIMD
#define VT_REF_WAVES_PER_SIMD 7
#endif
// kSched: the launch order comes from the frame's tile schedule (`sched`);
// the kernels without it carry none of its code.
template <class L, bool kStats, int kShade, bool kSched>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(kShade == kShadeRef ? VT_REF_WAVES_PER_SIMD : 7))) void k_render_dense(
    DenseWorld w, ViewConsts v, ShadeConsts sc, FrameOut out, uint32_t* sched, uint32_t log2_tw,
    uint32_t blocks_x) {
  __shared__ SchedBlock sb;
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
  if constexpr (kSched)
    bid = sched_begin(sched_geometry(sched, v.W, v.H, log2_tw, gridDim.x), bid, wave, sb);
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t bx = bid % blocks_x, by = bid / blocks_x;
  const uint32_t px = bx * (2u * tw) + (wave & 1u) * tw + (lane & (tw - 1u));
  const uint32_t py = by * (2u * th) + (wave >> 1) * th + (lane >> log2_tw);
  Counters cnt = {0, 0, 0, 0, 0, 0, 0, 0};
  GradAcc gacc = grad_acc();
  if (px < v.W && py < v.H)
    trace_pixel_dense<L, kStats, kShade>(w, v, sc, out, px, py, cnt, gacc);
  if (kStats) flush_counters(out, cnt);
  if constexpr (kShade == kShadeLambertGrad) {
    flush_albedo(sc, gacc);
    flush_grad(sc, gacc);
  }
  if constexpr (kSched) sched_end(sb, lane);
}

// Seven waves per SIMD (at most 72 VGPRs): the engine's own frame 0.373 vs
// 0.460 ms at eight (fewer spills around its many literal-loop iterations),
// the C1 world in this layout 0.823 vs 0.820 ms; five waves (90 VGPRs) were
// slower on both.  The free-region variants take six (80 VGPRs, 64 B of
// scratch per lane against 112 at seven): with the reach box and the blind
// prefix the C1 world in this layout measured 204.4 us at six against 212.5
// at seven, 206.6 at five and 225.9 at eight; the engine's frame (no
// free-region test) measures the same at six (275.7 against 275.4 us).
// kFree: the empty-space loops test the free-region map.
// kSched: as k_render_dense.
template <bool kStats, bool kFree, bool kSched>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kFree ? 6 : 7))) void k_render_brickmap(
    BrickmapWorld w, ViewConsts v, ShadeConsts sc, FrameOut out, uint32_t* sched,
    uint32_t log2_tw, uint32_t blocks_x) {
  __shared__ SchedBlock sb;
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
  if constexpr (kSched)
    bid = sched_begin(sched_geometry(sched, v.W, v.H, log2_tw, gridDim.x), bid, wave, sb);
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t lane = threadIdx.x & 63u;
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
    int blind = 0;
    if (kFree && !kStats && !line_meets_reach_bm(w, ex, ey, ez, dx, dy, dz, blind)) {
      // with the free-region test: this ray cannot reach a nonempty block (a
      // miss, no steps).  Without the test the check is left out: the
      // engine's own frame (caps 192/48, its camera inside the box, where
      // no ray can end) measured 0.304 ms with it against 0.274 without.
      g.data = 0u;
      g.hx = g.hy = g.hz = -1.0f;
      g.steps = 0;
    } else if (!kStats && w.occ)
      trace_brickmap_fast<false, kFree>(w, ex, ey, ez, dx, dy, dz, sc.max_primary, g, blind);
    else
      trace_general<kStats>(w, ex, ey, ez, dx, dy, dz, sc.max_primary, g, cnt.l1f, cnt.l0f);
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
      store_hit_primary(v, sc, out, (size_t)py * v.W + px, p);
      const uint32_t data = p.data, info0 = hit_info(p.face, p.sub, p.material, false);
      const int psteps = p.steps;
      GeneralHit s;
      if (!kStats && w.occ) {
        // the sun is the same for every lane: a uniform branch
        if (sc.sun[0] >= 0.0f && sc.sun[1] >= 0.0f && sc.sun[2] >= 0.0f)
          trace_brickmap_fast<true, kFree>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
                                           sc.max_shadow, s);
        else
          trace_brickmap_fast<false, kFree>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
                                            sc.max_shadow, s);
      } else {
        trace_general<kStats>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2], sc.max_shadow,
                              s, cnt.l1f, cnt.l0f);
      }
      shit = s.data != 0u;
      if (shit) {
        shx = s.hx;
        shy = s.hy;
        shz = s.hz;
      }
      ssteps = s.steps;
      store_hit_shadow_ref(v, sc, out, px, py, data, info0, psteps, shit, shx, shy, shz, sox, soy,
                           soz, ssteps, cnt);
    }
    if (!p.hit) {  // a hit's stores are made above
      GradAcc gacc = grad_acc();
      write_pixel<kShadeRef>(v, sc, out, px, py, p, shit, shx, shy, shz, sox, soy, soz, ssteps,
                             cnt, nullptr, gacc);
    }
  }
  if (kStats) flush_counters(out, cnt);
  if constexpr (kSched) sched_end(sb, lane);
}

static void grid_shape(const ViewConsts& v, TileShape tile, uint32_t& blocks_x, uint32_t& nblocks) {
  uint32_t tw = 1u << tile.log2_w, th = 64u >> tile.log2_w;
  blocks_x = (v.W + 2u * tw - 1u) / (2u * tw);
  uint32_t blocks_y = (v.H + 2u * th - 1u) / (2u * th);
  nblocks = blocks_x * blocks_y;
}


template <class L, bool kStats, bool kSched>
static void launch_dense_shade(const DenseWorld& w, const ViewConsts& v, const ShadeConsts& sc,
                               const FrameOut& out, uint32_t* sd, uint32_t log2_w,
                               uint32_t blocks_x, dim3 grid, hipStream_t stream) {
  switch (shade_mode(sc.flags)) {
    case kShadeLambertGrad:
      hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeLambertGrad, kSched>), grid, dim3(256),
                         0, stream, w, v, sc, out, sd, log2_w, blocks_x);
      break;
    case kShadeLambert:
      hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeLambert, kSched>), grid, dim3(256), 0,
                         stream, w, v, sc, out, sd, log2_w, blocks_x);
      break;
    default:
      hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeRef, kSched>), grid, dim3(256), 0,
                         stream, w, v, sc, out, sd, log2_w, blocks_x);
  }
}

// The traversal variant of a dense launch (kSched: as k_render_dense).
template <bool kSched>
static void launch_dense_select(const DenseWorld& w, uint32_t traversal, const ViewConsts& v,
                                const ShadeConsts& sc, const FrameOut& out, uint32_t* sd,
                                uint32_t lw, uint32_t blocks_x, dim3 grid, hipStream_t stream) {
  const bool stats = out.stats != nullptr;
  if (traversal == VT_TRAVERSAL_BITMASK) {
    if (stats)
      launch_dense_shade<BitmaskLookup, true, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
    else
      launch_dense_shade<BitmaskLookup, false, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
  } else if (traversal == VT_TRAVERSAL_BITMASK8) {
    if (stats)
      launch_dense_shade<BitmaskLookup8, true, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
    else
      launch_dense_shade<BitmaskLookup8, false, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
  } else {
    const bool long_caps = sc.max_primary >= kLongCap || sc.max_shadow >= kLongCap;
    if (stats && w.free_bytes)  // statistics of the steps executed
      launch_dense_shade<FlatLookupFree, true, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
    else if (stats)
      launch_dense_shade<FlatLookup, true, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
    else if (w.free_bytes)
      launch_dense_shade<FlatLookupFree, false, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
    else if (long_caps)
      launch_dense_shade<FlatLookupLong, false, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
    else
      launch_dense_shade<FlatLookup, false, kSched>(w, v, sc, out, sd, lw, blocks_x, grid, stream);
  }
}

// The tile schedule is taken for frames with a buffer and a step cap of at
// least kLongCap.  With the shaders' own caps (192 / 48) the blocks' costs
// are flat and the costliest-first order measured slower: the engine's frame
// +8.6% at 1280x720, +19% at 1920x1080, the 512^3 world +4% at 1280x720;
// with caps of 2048 it gains 7-25% on both worlds at 1280x720..2560x1440.
static bool sched_taken(const uint32_t* schedule, const ShadeConsts& sc) {
  return schedule && (sc.max_primary >= kLongCap || sc.max_shadow >= kLongCap);
}

hipError_t launch_render_dense(const DenseWorld& w_in, uint32_t traversal, const ViewConsts& v,
                               const ShadeConsts& sc_in, const FrameOut& out, uint32_t* schedule,
                               TileShape tile, uint32_t compute_units, hipStream_t stream) {
  (void)compute_units;
  DenseWorld w = w_in;
  if (out.steps) w.free_bytes = 0;  // step counts are stored: no early end of rays
  // nor where they are stored, or counted as traceMap's (statistics without
  // the free-region test), for the reach box
  if (out.steps || (out.stats && !w.free_bytes)) reach_everywhere(w.reach_lo, w.reach_hi);
  ShadeConsts sc = sc_in;
  sc.shadow = make_ray_consts(sc.sun[0], sc.sun[1], sc.sun[2], w.mx, w.my, w.mz);
  uint32_t blocks_x, nblocks;
  grid_shape(v, tile, blocks_x, nblocks);
  if (schedule && nblocks > VT_SCHEDULE_BLOCKS(v.W, v.H)) return hipErrorInvalidValue;
  const dim3 grid(nblocks);
  if (sched_taken(schedule, sc))
    launch_dense_select<true>(w, traversal, v, sc, out, schedule, tile.log2_w, blocks_x, grid,
                              stream);
  else
    launch_dense_select<false>(w, traversal, v, sc, out, nullptr, tile.log2_w, blocks_x, grid,
                               stream);
  return hipGetLastError();
}

template <bool kStats, bool kFree>
static void launch_brickmap(bool sched, const BrickmapWorld& w, const ViewConsts& v,
                            const ShadeConsts& sc, const FrameOut& out, uint32_t* schedule,
                            uint32_t lw, uint32_t blocks_x, dim3 grid, hipStream_t stream) {
  if (sched)
    hipLaunchKernelGGL((k_render_brickmap<kStats, kFree, true>), grid, dim3(256), 0, stream, w, v,
                       sc, out, schedule, lw, blocks_x);
  else
    hipLaunchKernelGGL((k_render_brickmap<kStats, kFree, false>), grid, dim3(256), 0, stream, w,
                       v, sc, out, nullptr, lw, blocks_x);
}

hipError_t launch_render_brickmap(const BrickmapWorld& w_in, const ViewConsts& v,
                                  const ShadeConsts& sc, const FrameOut& out, uint32_t* schedule,
                                  TileShape tile, uint32_t compute_units, hipStream_t stream) {
  (void)compute_units;
  BrickmapWorld w = w_in;
  // step counts stored, or statistics (of traceMap's steps on this layout):
  // no early end of rays
  if (out.steps || out.stats) {
    w.free_bytes = 0;
    reach_everywhere(w.reach_lo, w.reach_hi);
  }
  uint32_t blocks_x, nblocks;
  grid_shape(v, tile, blocks_x, nblocks);
  if (schedule && nblocks > VT_SCHEDULE_BLOCKS(v.W, v.H)) return hipErrorInvalidValue;
  const dim3 grid(nblocks);
  const bool sched = sched_taken(schedule, sc);
  const uint32_t lw = tile.log2_w;
  if (out.stats)
    launch_brickmap<true, false>(sched, w, v, sc, out, schedule, lw, blocks_x, grid, stream);
  else if (w.occ && w.free_bytes)
    launch_brickmap<false, true>(sched, w, v, sc, out, schedule, lw, blocks_x, grid, stream);
  else
    launch_brickmap<false, false>(sched, w, v, sc, out, schedule, lw, blocks_x, grid, stream);
  return hipGetLastError();
}

}  // namespace vt
