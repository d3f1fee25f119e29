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
// Shadow rays come in two kinds: most of those not lit are occluded after a
// step or two, most lit ones run to the edge of the grid (hundreds of steps).
// In a wave that holds both, the occluded lanes idle while the lit ones run.
// So (reference composite, flat lookup) each lane runs its shadow ray for a
// first segment of one empty-space iteration; the rays still running are
// then parked in LDS, packed, and dealt out to the block's waves from the
// first on, 64 to a wave, which finish them -- the block's waves without a
// parked ray end there.  Same steps and results per ray; one block-wide
// barrier on each side of the packing.  (A second kernel over the long rays
// of the whole frame, through a job list in memory, was built and measured
// slower than the single kernel once the XCD strips were in.)
#include "vt_pixel.hpp"

// Frames with at least this many 256-thread blocks per CU, and a shadow cap
// of at least VT_PACK_MIN_SHADOW_CAP steps, take the packed shadow rays
// (kPack of k_render_dense and k_render_brickmap).  The packing costs two
// block-wide barriers: it pays when the frame fills every CU's block slots
// (1080p faster, 512x512 views slower) and when shadow rays can run long --
// the imbalance it removes is bounded by the shadow cap (the engine's own
// frame, cap 48, measured 16% slower with it in the brickmap kernel).
#ifndef VT_PACK_BLOCKS_PER_CU
#define VT_PACK_BLOCKS_PER_CU 8u
#endif
#ifndef VT_PACK_MIN_SHADOW_CAP
#define VT_PACK_MIN_SHADOW_CAP 256
#endif

namespace vt {

// amdgpu_waves_per_eu: the reference composite at most 64 VGPRs (eight waves
// per SIMD hide the gathers of the traversal loop; the spills it costs lie
// outside it; 1080p 0.785 vs 0.804 ms at seven), the Lambert modes at most
// 72 (seven waves: C3's Lambert forward 0.117 vs 0.120 ms per view).
// kPack: the block packs its long shadow rays (reference composite, flat
// lookup, no statistics); the launcher picks it for frames that fill the GPU.
template <class L, bool kStats, int kShade, bool kPack = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(kShade == kShadeRef ? 8 : 7))) void k_render_dense(
    DenseWorld w, ViewConsts v, ShadeConsts sc, FrameOut out, uint32_t log2_tw,
    uint32_t blocks_x) {
  const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t bx = bid % blocks_x, by = bid / blocks_x;
  const uint32_t px = bx * (2u * tw) + (wave & 1u) * tw + (lane & (tw - 1u));
  const uint32_t py = by * (2u * th) + (wave >> 1) * th + (lane >> log2_tw);
  Counters cnt = {0, 0, 0, 0, 0, 0, 0, 0};
  if constexpr (kPack && kShade == kShadeRef && L::kPrefetch && !kStats) {
    // every thread of the block reaches both barriers: no early exit above
    __shared__ ShadowJob jobs[256];
    __shared__ uint32_t wave_jobs[4];
    ShadowJob job;
    bool surv = false;
    if (px < v.W && py < v.H) surv = trace_pixel_dense_split<L>(w, v, sc, out, px, py, cnt, job);
    const unsigned long long bal = __ballot(surv);
    if (lane == 0) wave_jobs[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
      const uint32_t n = wave_jobs[i];
      base += i < wave ? n : 0u;
      total += n;
    }
    if (surv) jobs[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = job;
    __syncthreads();
    if (threadIdx.x < total) finish_shadow_job<L>(w, v, sc, out, jobs[threadIdx.x], cnt);
    return;
  }
  GradAcc gacc = grad_acc();
  if (px < v.W && py < v.H)
    trace_pixel_dense<L, kStats, kShade>(w, v, sc, out, px, py, cnt, gacc);
  if (kStats) flush_counters(out, cnt);
  if constexpr (kShade == kShadeLambertGrad) {
    flush_albedo(sc, gacc);
    flush_grad(sc, gacc);
  }
}

// Seven waves per SIMD (at most 72 VGPRs): the engine's own frame 0.373 vs
// 0.460 ms at eight (fewer spills around its many literal-loop iterations),
// the C1 world in this layout 0.823 vs 0.820 ms; five waves (90 VGPRs) were
// slower on both.
// kPack: the block packs its long shadow rays after a first segment and
// finishes them 64 to a wave, as k_render_dense does (worlds with occupancy
// bytes, no statistics).
template <bool kStats, bool kPack = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_render_brickmap(
    BrickmapWorld w, ViewConsts v, ShadeConsts sc, FrameOut out, uint32_t log2_tw,
    uint32_t blocks_x) {
  const uint32_t bid = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t bx = bid % blocks_x, by = bid / blocks_x;
  const uint32_t px = bx * (2u * tw) + (wave & 1u) * tw + (lane & (tw - 1u));
  const uint32_t py = by * (2u * th) + (wave >> 1) * th + (lane >> log2_tw);
  const bool active = px < v.W && py < v.H;
  Counters cnt = {0, 0, 0, 0, 0, 0, 0, 0};
  if constexpr (kPack && !kStats) {
    // (launched only for worlds with occupancy bytes) every thread of the
    // block reaches both barriers
    __shared__ ShadowJob jobs[256];
    __shared__ uint32_t wave_jobs[4];
    ShadowJob job;
    bool surv = false;
    const bool pos = sc.sun[0] >= 0.0f && sc.sun[1] >= 0.0f && sc.sun[2] >= 0.0f;
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
      trace_brickmap_fast(w, ex, ey, ez, dx, dy, dz, sc.max_primary, g);
      Primary p;
      p.hit = g.data != 0u;
      p.steps = g.steps;
      if (p.hit) {
        p.face = g.face;
        p.voxel = (uint32_t)g.bx + w.dim * ((uint32_t)g.by + w.dim * (uint32_t)g.bz);
        p.data = g.data;
        p.material = bm_block(w, g.bx, g.by, g.bz) & 255u;
        p.sub = g.sub;
        p.hx = g.hx;
        p.hy = g.hy;
        p.hz = g.hz;
        float sox, soy, soz;
        shadow_origin(p, sc.flags, sox, soy, soz);
        store_hit_primary(v, sc, out, (size_t)py * v.W + px, p);
        const uint32_t data = p.data, info0 = hit_info(p.face, p.sub, p.material, false);
        GeneralHit sh;
        BrickPark park;
        if (pos)
          trace_brickmap_fast<true, true>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
                                          sc.max_shadow, sh, &park);
        else
          trace_brickmap_fast<false, true>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
                                           sc.max_shadow, sh, &park);
        if (sh.steps < 0) {
          surv = true;
          job.wx = park.wx;
          job.wy = park.wy;
          job.wz = park.wz;
          job.c = park.c;
          job.steps = park.steps;
          job.pxy = px | (py << 16);
          job.data = data;
          job.info0 = info0;
          job.psteps = p.steps;
          job.sox = sox;
          job.soy = soy;
          job.soz = soz;
        } else {
          const bool shit = sh.data != 0u;
          store_hit_shadow_ref(v, sc, out, px, py, data, info0, p.steps, shit,
                               shit ? sh.hx : -1.0f, shit ? sh.hy : -1.0f, shit ? sh.hz : -1.0f,
                               sox, soy, soz, sh.steps, cnt);
        }
      } else {
        GradAcc gacc = grad_acc();
        write_pixel<kShadeRef>(v, sc, out, px, py, p, false, -1.0f, -1.0f, -1.0f, 0.f, 0.f, 0.f, 0,
                               cnt, nullptr, gacc);
      }
    }
    const unsigned long long bal = __ballot(surv);
    if (lane == 0) wave_jobs[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; i++) {
      const uint32_t n = wave_jobs[i];
      base += i < wave ? n : 0u;
      total += n;
    }
    if (surv) jobs[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = job;
    __syncthreads();
    if (threadIdx.x < total) {
      const ShadowJob& jb = jobs[threadIdx.x];
      BrickPark from;
      from.wx = jb.wx;
      from.wy = jb.wy;
      from.wz = jb.wz;
      from.c = jb.c;
      from.steps = jb.steps;
      GeneralHit sh;
      if (pos)
        trace_brickmap_fast<true, false, true>(w, 0.f, 0.f, 0.f, sc.sun[0], sc.sun[1], sc.sun[2],
                                               sc.max_shadow, sh, nullptr, &from);
      else
        trace_brickmap_fast<false, false, true>(w, 0.f, 0.f, 0.f, sc.sun[0], sc.sun[1],
                                                sc.sun[2], sc.max_shadow, sh, nullptr, &from);
      const bool shit = sh.data != 0u;
      store_hit_shadow_ref(v, sc, out, jb.pxy & 0xFFFFu, jb.pxy >> 16, jb.data, jb.info0,
                           jb.psteps, shit, shit ? sh.hx : -1.0f, shit ? sh.hy : -1.0f,
                           shit ? sh.hz : -1.0f, jb.sox, jb.soy, jb.soz, sh.steps, cnt);
    }
    return;
  }
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
    if (!kStats && w.occ)
      trace_brickmap_fast(w, ex, ey, ez, dx, dy, dz, sc.max_primary, g);
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
          trace_brickmap_fast<true>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
                                    sc.max_shadow, s);
        else
          trace_brickmap_fast<false>(w, sox, soy, soz, sc.sun[0], sc.sun[1], sc.sun[2],
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
}

static void grid_shape(const ViewConsts& v, TileShape tile, uint32_t& blocks_x, uint32_t& nblocks) {
  uint32_t tw = 1u << tile.log2_w, th = 64u >> tile.log2_w;
  blocks_x = (v.W + 2u * tw - 1u) / (2u * tw);
  uint32_t blocks_y = (v.H + 2u * th - 1u) / (2u * th);
  nblocks = blocks_x * blocks_y;
}

template <class L, bool kStats>
static void launch_dense_shade(const DenseWorld& w, const ViewConsts& v, const ShadeConsts& sc,
                               const FrameOut& out, uint32_t log2_w, uint32_t blocks_x, dim3 grid,
                               bool pack, hipStream_t stream) {
  switch (shade_mode(sc.flags)) {
    case kShadeLambertGrad:
      hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeLambertGrad>), grid, dim3(256), 0,
                         stream, w, v, sc, out, log2_w, blocks_x);
      break;
    case kShadeLambert:
      hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeLambert>), grid, dim3(256), 0, stream, w,
                         v, sc, out, log2_w, blocks_x);
      break;
    default:
      if constexpr (L::kPrefetch && !kStats) {
        if (pack) {
          hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeRef, true>), grid, dim3(256), 0,
                             stream, w, v, sc, out, log2_w, blocks_x);
          break;
        }
      }
      hipLaunchKernelGGL((k_render_dense<L, kStats, kShadeRef>), grid, dim3(256), 0, stream, w, v,
                         sc, out, log2_w, blocks_x);
  }
}

hipError_t launch_render_dense(const DenseWorld& w, uint32_t traversal, const ViewConsts& v,
                               const ShadeConsts& sc_in, const FrameOut& out, TileShape tile,
                               uint32_t compute_units, hipStream_t stream) {
  ShadeConsts sc = sc_in;
  sc.shadow = make_ray_consts(sc.sun[0], sc.sun[1], sc.sun[2], w.mx, w.my, w.mz);
  uint32_t blocks_x, nblocks;
  grid_shape(v, tile, blocks_x, nblocks);
  const bool pack = nblocks >= VT_PACK_BLOCKS_PER_CU * compute_units &&
                    sc.max_shadow >= VT_PACK_MIN_SHADOW_CAP;
  const dim3 grid(nblocks);
  const bool stats = out.stats != nullptr;
  if (traversal == VT_TRAVERSAL_BITMASK) {
    if (stats)
      launch_dense_shade<BitmaskLookup, true>(w, v, sc, out, tile.log2_w, blocks_x, grid, pack, stream);
    else
      launch_dense_shade<BitmaskLookup, false>(w, v, sc, out, tile.log2_w, blocks_x, grid, pack, stream);
  } else if (traversal == VT_TRAVERSAL_BITMASK8) {
    if (stats)
      launch_dense_shade<BitmaskLookup8, true>(w, v, sc, out, tile.log2_w, blocks_x, grid, pack, stream);
    else
      launch_dense_shade<BitmaskLookup8, false>(w, v, sc, out, tile.log2_w, blocks_x, grid, pack, stream);
  } else {
    if (stats)
      launch_dense_shade<FlatLookup, true>(w, v, sc, out, tile.log2_w, blocks_x, grid, pack, stream);
    else
      launch_dense_shade<FlatLookup, false>(w, v, sc, out, tile.log2_w, blocks_x, grid, pack, stream);
  }
  return hipGetLastError();
}

hipError_t launch_render_brickmap(const BrickmapWorld& w, const ViewConsts& v,
                                  const ShadeConsts& sc, const FrameOut& out, TileShape tile,
                                  uint32_t compute_units, hipStream_t stream) {
  uint32_t blocks_x, nblocks;
  grid_shape(v, tile, blocks_x, nblocks);
  dim3 grid(nblocks), block(256);
  if (out.stats)
    hipLaunchKernelGGL((k_render_brickmap<true>), grid, block, 0, stream, w, v, sc, out,
                       tile.log2_w, blocks_x);
  else if (w.occ && nblocks >= VT_PACK_BLOCKS_PER_CU * compute_units &&
           sc.max_shadow >= VT_PACK_MIN_SHADOW_CAP)
    hipLaunchKernelGGL((k_render_brickmap<false, true>), grid, block, 0, stream, w, v, sc, out,
                       tile.log2_w, blocks_x);
  else
    hipLaunchKernelGGL((k_render_brickmap<false>), grid, block, 0, stream, w, v, sc, out,
                       tile.log2_w, blocks_x);
  return hipGetLastError();
}

}  // namespace vt
