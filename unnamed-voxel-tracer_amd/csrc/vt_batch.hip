// vt_batch.hip -- many views of uneven resolution in one persistent launch.
//
// Work unit: a 64-pixel tile (one wave).  The tiles of all views are numbered
// view by view (tile_base = prefix sum of tiles per view) and split into 8
// contiguous ranges, one per XCD: a block dealt to XCD x (blocks are dealt
// round-robin, x = blockIdx % 8) pulls `chunk` tiles at a time from queue x
// with one atomic, so neighbouring tiles -- which touch neighbouring voxels --
// stay in one XCD's L2; when its queue runs dry a wave steals from the other
// queues in order.  Every wave reaches the end of all 8 queues, so the grid
// drains.  The per-pixel work is trace_pixel_dense, shared with vt_render.
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "vt_pixel.hpp"

namespace vt {

struct BatchView {
  ViewConsts v;
  FrameOut out;
  uint32_t tiles_x;
  uint32_t tile_base;
};

constexpr uint32_t kQueues = 8;

__global__ void k_batch_reset(uint32_t* __restrict__ queues) {
  if (threadIdx.x < kQueues) queues[threadIdx.x] = 0u;
}

// 7 waves per SIMD (72 VGPRs): the view record and queue state push the
// kernel to ~94 VGPRs (5 waves); capping it spills values that live outside
// the traversal loops and hides the gathers' latency far better (C2: 19.5 ->
// 16.3 ms per 64-view batch at 8 waves).  With the free-region test seven
// waves measure 3.81 / 3.84 ms against 4.01 / 3.91 at eight (fewer spills:
// 132 vs 176 B per lane).
#ifndef VT_BATCH_WAVES_PER_SIMD
#define VT_BATCH_WAVES_PER_SIMD 7
#endif
template <class L, bool kStats, int kShade>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VT_BATCH_WAVES_PER_SIMD))) void k_render_batch(DenseWorld w, ShadeConsts sc,
                                                      const BatchView* __restrict__ views,
                                                      uint32_t n_views, uint32_t total_tiles,
                                                      uint32_t* __restrict__ queues,
                                                      uint32_t chunk, uint32_t log2_tw,
                                                      unsigned long long* stats) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  const uint32_t home = blockIdx.x % kQueues;
  Counters cnt = {0, 0, 0, 0, 0, 0, 0, 0};
  GradAcc gacc = grad_acc();
  for (uint32_t k = 0; k < kQueues; k++) {
    const uint32_t q = (home + k) % kQueues;
    const uint32_t lo = (uint32_t)(((uint64_t)total_tiles * q) / kQueues);
    const uint32_t hi = (uint32_t)(((uint64_t)total_tiles * (q + 1)) / kQueues);
    for (;;) {
      uint32_t got = 0;
      if (lane == 0) got = atomicAdd(queues + q, chunk);
      got = __builtin_amdgcn_readfirstlane(__shfl(got, 0, 64));
      if (got >= hi - lo) break;
      const uint32_t t1 = lo + min(got + chunk, hi - lo);
      for (uint32_t t = lo + got; t < t1; t++) {
        // the view owning tile t: last view whose tile_base <= t
        uint32_t a = 0, b = n_views;
        while (b - a > 1u) {
          const uint32_t m = (a + b) >> 1;
          if (views[m].tile_base <= t) a = m;
          else b = m;
        }
        a = __builtin_amdgcn_readfirstlane(a);
        const BatchView& bv = views[a];
        const uint32_t local = t - bv.tile_base;
        const uint32_t ty = local / bv.tiles_x, tx = local - ty * bv.tiles_x;
        const uint32_t px = tx * tw + (lane & (tw - 1u));
        const uint32_t py = ty * th + (lane >> log2_tw);
        const ViewConsts v = bv.v;
        if (px < v.W && py < v.H)
          trace_pixel_dense<L, kStats, kShade>(w, v, sc, bv.out, px, py, cnt, gacc);
        if constexpr (kShade == kShadeLambertGrad) flush_albedo(sc, gacc);  // one pixel per lane
      }
    }
  }
  if (kStats) {
    FrameOut o;
    o.stats = stats;
    flush_counters(o, cnt);
  }
  if constexpr (kShade == kShadeLambertGrad) flush_grad(sc, gacc);
}

// ---- the no-allocation batch (vt_batch_render_noalloc) ----------------------
// Its view table and queue counters live in a file-scope device array (one
// copy per GPU, held until the process ends) instead of memory the library
// allocates.  Host code never passes the array's name to the device: its
// address on each device comes from hipGetSymbolAddress, taken once per
// device at vt_create (noalloc_init) with that device current.
struct NoallocScratch {
  BatchView views[VT_NOALLOC_MAX_VIEWS];
  uint32_t queues[kQueues];
};

static __device__ __align__(256) NoallocScratch scratch;

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static NoallocScratch* g_scratch_ptr[kMaxDevices];  // per device ordinal, null until set

int noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(VT_ERR_NO_DEVICE, "vt_create: device %d beyond the %d this library tracks",
                device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return VT_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                    "vt_create: address of the no-allocation scratch");
  }
  g_scratch_ptr[device] = static_cast<NoallocScratch*>(p);
  return VT_OK;
}

static NoallocScratch* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

// Views handed to the device as kernel arguments (copied by the launch
// itself, so the host table may go away at once), written into the scratch
// view table: up to kStageViews per launch.
constexpr uint32_t kStageViews = 8;
struct StagePack {
  BatchView v[kStageViews];
  uint32_t first, count;
};

__global__ void k_stage_views(BatchView* __restrict__ dst, StagePack pack) {
  const uint32_t i = threadIdx.x;
  if (i < pack.count) dst[pack.first + i] = pack.v[i];
}

}  // namespace vt

struct vt_batch {
  vt_ctx* ctx;
  const vt_world* world;
  uint32_t n_views, total_tiles, nblocks;
  vt_tuning tuning;
  vt::ShadeConsts sc;
  vt::BatchView* d_views;
  uint32_t* d_queues;
  unsigned long long* stats;
  bool steps_out;  // a view stores step counts: rays are not ended early
  hipStream_t cap_stream;
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace {

using vt::fail;
using vt::hip_fail;

// Queue reset + persistent launch over a device view table.
// steps_out: some view stores step counts, which the free-region map's early
// end of a ray would change, so the map is not read.
hipError_t launch_batch(const vt_world* world, const vt::ShadeConsts& sc_in,
                        const vt::BatchView* d_views, uint32_t n_views, uint32_t total_tiles,
                        uint32_t* d_queues, const vt_tuning& tuning, uint32_t nblocks,
                        unsigned long long* stats, bool steps_out, hipStream_t s) {
  vt::DenseWorld w = vt::dense_view(world);
  if (steps_out || stats || !vt::use_free_map(world, sc_in)) w.free_bytes = 0;
  // rays end early at the reach box unless their step counts are observed
  if (steps_out || stats) vt::reach_everywhere(w.reach_lo, w.reach_hi);
  vt::ShadeConsts sc = sc_in;
  sc.shadow = vt::make_ray_consts(sc.sun[0], sc.sun[1], sc.sun[2], w.mx, w.my, w.mz);
  hipLaunchKernelGGL(vt::k_batch_reset, dim3(1), dim3(64), 0, s, d_queues);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const dim3 grid(nblocks), block(256);
  const uint32_t tile = tuning.tile, chunk = tuning.chunk;
#define VT_BATCH_LAUNCH(L, S, M)                                                           \
  hipLaunchKernelGGL((vt::k_render_batch<vt::L, S, vt::M>), grid, block, 0, s, w, sc, d_views, \
                     n_views, total_tiles, d_queues, chunk, tile, stats)
#define VT_BATCH_SHADE(L, S)                                                            \
  switch (vt::shade_mode(sc.flags)) {                                                   \
    case vt::kShadeLambertGrad: VT_BATCH_LAUNCH(L, S, kShadeLambertGrad); break;        \
    case vt::kShadeLambert: VT_BATCH_LAUNCH(L, S, kShadeLambert); break;                \
    default: VT_BATCH_LAUNCH(L, S, kShadeRef);                                          \
  }
  if (tuning.traversal == VT_TRAVERSAL_BITMASK) {
    if (stats) VT_BATCH_SHADE(BitmaskLookup, true)
    else VT_BATCH_SHADE(BitmaskLookup, false)
  } else if (tuning.traversal == VT_TRAVERSAL_BITMASK8) {
    if (stats) VT_BATCH_SHADE(BitmaskLookup8, true)
    else VT_BATCH_SHADE(BitmaskLookup8, false)
  } else {
    if (stats) VT_BATCH_SHADE(FlatLookup, true)
    else if (w.free_bytes) VT_BATCH_SHADE(FlatLookupFree, false)
    else if (sc.max_primary >= vt::kLongCap || sc.max_shadow >= vt::kLongCap)
      VT_BATCH_SHADE(FlatLookupLong, false)
    else VT_BATCH_SHADE(FlatLookup, false)
  }
#undef VT_BATCH_SHADE
#undef VT_BATCH_LAUNCH
  return hipGetLastError();
}

hipError_t enqueue(const vt_batch* b, hipStream_t s) {
  return launch_batch(b->world, b->sc, b->d_views, b->n_views, b->total_tiles, b->d_queues,
                      b->tuning, b->nblocks, b->stats, b->steps_out, s);
}

// Validation shared by vt_batch_create and vt_batch_render_noalloc, and the
// view record of view i: *tiles_out receives the total tile count.
int check_batch(vt_ctx* ctx, const vt_world* world, const vt_view* views, uint32_t n_views,
                const vt_render_params* params, const vt_tuning* tuning,
                unsigned long long* stats, uint32_t max_views, const char* who,
                vt_render_params* p_out) {
  if (!ctx || !world || !views || !params || !tuning)
    return fail(VT_ERR_NULL_POINTER, "%s: ctx, world, views, params, tuning required", who);
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "%s: world from another ctx", who);
  if (world->kind != WorldKind::Dense)
    return fail(VT_ERR_UNSUPPORTED, "%s: batched rendering needs a dense world", who);
  if (n_views == 0 || n_views > max_views)
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: n_views %u outside [1, %u]", who, n_views,
                max_views);
  const vt_tuning t = *tuning;
  if (t.tile < VT_TILE_8X8 || t.tile > VT_TILE_32X2 || t.chunk < 1 || t.chunk > 64 ||
      t.traversal > VT_TRAVERSAL_BITMASK8 || t.blocks_per_cu < 1 || t.blocks_per_cu > 16)
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: tuning out of range", who);
  if (stats && !vt::aligned(stats, 8)) return fail(VT_ERR_MISALIGNED, "%s: stats", who);
  vt_render_params p = *params;
  p.traversal = t.traversal;
  int st = vt::check_params(p, world, who);
  if (st != VT_OK) return st;
  for (uint32_t i = 0; i < n_views; i++)
    if ((st = vt::check_image_size(views[i].width, views[i].height, who)) != VT_OK ||
        (st = vt::check_frame(&views[i].frame, p.flags, who)) != VT_OK)
      return st;
  *p_out = p;
  return VT_OK;
}

// View record of views[i] whose tiles start at *tiles (advanced past them).
int view_record(const vt_view& v, uint32_t log2_tw, uint64_t* tiles, vt::BatchView* b,
                const char* who) {
  const uint32_t tw = 1u << log2_tw, th = 64u >> log2_tw;
  std::memset(b, 0, sizeof *b);
  b->v = vt::view_consts(&v.camera, v.width, v.height);
  b->out = vt::frame_out(&v.frame);
  b->out.stats = nullptr;
  b->tiles_x = (v.width + tw - 1) / tw;
  b->tile_base = (uint32_t)*tiles;
  *tiles += (uint64_t)b->tiles_x * ((v.height + th - 1) / th);
  if (*tiles >= (1ull << 31))
    return fail(VT_ERR_INVALID_ARGUMENT, "%s: more than 2^31 tiles", who);
  return VT_OK;
}

void release(vt_batch* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  vt::dev_free(b->d_views);
  vt::dev_free(b->d_queues);
  delete b;
}

}  // namespace

extern "C" {

int vt_batch_create(vt_ctx* ctx, const vt_world* world, const vt_view* views, uint32_t n_views,
                    const vt_render_params* params, const vt_tuning* tuning,
                    unsigned long long* stats, vt_batch** out) {
  if (!out) return fail(VT_ERR_NULL_POINTER, "vt_batch_create: out is NULL");
  *out = nullptr;
  vt_render_params p;
  int st = check_batch(ctx, world, views, n_views, params, tuning, stats, 65536,
                       "vt_batch_create", &p);
  if (st != VT_OK) return st;
  const vt_tuning t = *tuning;
  std::vector<vt::BatchView> table(n_views);
  uint64_t tiles = 0;
  for (uint32_t i = 0; i < n_views; i++)
    if ((st = view_record(views[i], t.tile, &tiles, &table[i], "vt_batch_create")) != VT_OK)
      return st;
  VT_HIP(hipSetDevice(ctx->device));
  vt_batch* b = new (std::nothrow) vt_batch;
  if (!b) return fail(VT_ERR_OUT_OF_MEMORY, "vt_batch_create: host allocation");
  std::memset(b, 0, sizeof *b);
  b->ctx = ctx;
  b->world = world;
  b->n_views = n_views;
  b->total_tiles = (uint32_t)tiles;
  b->tuning = t;
  b->sc = vt::shade_consts(p);
  b->stats = stats;
  for (uint32_t i = 0; i < n_views; i++) b->steps_out |= views[i].frame.steps != nullptr;
  b->nblocks = (uint32_t)ctx->compute_units * t.blocks_per_cu;
  hipError_t e = hipMalloc((void**)&b->d_views, sizeof(vt::BatchView) * n_views);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_queues, sizeof(uint32_t) * vt::kQueues);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_views, table.data(), sizeof(vt::BatchView) * n_views,
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    release(b);
    return hip_fail(e, "vt_batch_create: device table");
  }
  *out = b;
  return VT_OK;
}

int vt_batch_launch(vt_batch* batch, void* stream) {
  if (!batch) return fail(VT_ERR_NULL_POINTER, "vt_batch_launch: batch is NULL");
  VT_HIP(hipSetDevice(batch->ctx->device));
  hipError_t e = enqueue(batch, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "vt_batch_launch");
  return VT_OK;
}

int vt_batch_capture_graph(vt_batch* b) {
  if (!b) return fail(VT_ERR_NULL_POINTER, "vt_batch_capture_graph: batch is NULL");
  VT_HIP(hipSetDevice(b->ctx->device));
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  if (!b->cap_stream) VT_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
  VT_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue(b, b->cap_stream);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->cap_stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(e != hipSuccess ? e : e2, "vt_batch_capture_graph: capture");
  }
  b->graph = g;
  VT_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return VT_OK;
}

int vt_batch_graph_launch(vt_batch* b, void* stream) {
  if (!b) return fail(VT_ERR_NULL_POINTER, "vt_batch_graph_launch: batch is NULL");
  if (!b->exec) return fail(VT_ERR_INVALID_ARGUMENT, "vt_batch_graph_launch: no captured graph");
  VT_HIP(hipSetDevice(b->ctx->device));
  VT_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return VT_OK;
}

int vt_batch_destroy(vt_batch* b) {
  if (!b) return fail(VT_ERR_NULL_POINTER, "vt_batch_destroy: batch is NULL");
  (void)hipSetDevice(b->ctx->device);
  release(b);
  return VT_OK;
}

int vt_batch_autotune(vt_ctx* ctx, const vt_world* world, const vt_view* views,
                      uint32_t n_views, const vt_render_params* params, int reps,
                      vt_tuning* best, float* best_ms) {
  if (!best || !best_ms) return fail(VT_ERR_NULL_POINTER, "vt_batch_autotune: outputs required");
  if (!ctx) return fail(VT_ERR_NULL_POINTER, "vt_batch_autotune: ctx is NULL");
  if (reps < 1 || reps > 1000)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_batch_autotune: reps %d outside [1, 1000]", reps);
  if (params && (params->flags & VT_FLAG_GRAD_L2))
    return fail(VT_ERR_INVALID_ARGUMENT,
                "vt_batch_autotune: VT_FLAG_GRAD_L2 accumulates into the caller's gradients on "
                "every timed launch; tune with VT_FLAG_LAMBERT alone (same kernel shape)");
  VT_HIP(hipSetDevice(ctx->device));
  hipStream_t s;
  VT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best_t = 1e30f;
  vt_tuning best_cfg = {VT_TILE_8X8, 1, VT_TRAVERSAL_FLAT, 8};
  int st = VT_OK;
  const uint32_t tiles[] = {VT_TILE_8X8, VT_TILE_16X4, VT_TILE_32X2};
  const uint32_t chunks[] = {1, 2, 4, 8};
  const uint32_t travs[] = {VT_TRAVERSAL_FLAT, VT_TRAVERSAL_BITMASK, VT_TRAVERSAL_BITMASK8};
  for (uint32_t tr : travs)
    for (uint32_t ti : tiles)
      for (uint32_t ch : chunks) {
        vt_tuning cfg = {ti, ch, tr, 8};
        vt_batch* b = nullptr;
        st = vt_batch_create(ctx, world, views, n_views, params, &cfg, nullptr, &b);
        if (st != VT_OK) goto done;
        st = vt_batch_launch(b, s);  // warm-up
        if (st == VT_OK && hipEventRecord(e0, s) != hipSuccess) st = VT_ERR_DEVICE;
        for (int r = 0; st == VT_OK && r < reps; r++) st = vt_batch_launch(b, s);
        if (st == VT_OK && hipEventRecord(e1, s) != hipSuccess) st = VT_ERR_DEVICE;
        if (st == VT_OK && hipEventSynchronize(e1) != hipSuccess) st = VT_ERR_DEVICE;
        float ms = 0.f;
        if (st == VT_OK) (void)hipEventElapsedTime(&ms, e0, e1);
        vt_batch_destroy(b);
        if (st != VT_OK) goto done;
        ms /= (float)reps;
        if (ms < best_t) {
          best_t = ms;
          best_cfg = cfg;
        }
      }
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (st != VT_OK) return st == VT_ERR_DEVICE ? fail(st, "vt_batch_autotune: timing") : st;
  *best = best_cfg;
  *best_ms = best_t;
  return VT_OK;
}

int vt_batch_render_noalloc(vt_ctx* ctx, const vt_world* world, const vt_view* views,
                            uint32_t n_views, const vt_render_params* params,
                            const vt_tuning* tuning, unsigned long long* stats, void* stream) {
  vt_render_params p;
  int st = check_batch(ctx, world, views, n_views, params, tuning, stats, VT_NOALLOC_MAX_VIEWS,
                       "vt_batch_render_noalloc", &p);
  if (st != VT_OK) return st;
  VT_HIP(hipSetDevice(ctx->device));
  vt::NoallocScratch* sp = vt::noalloc_scratch(ctx->device);
  if (!sp) return fail(VT_ERR_DEVICE, "vt_batch_render_noalloc: no scratch address for device %d",
                       ctx->device);
  const hipStream_t s = (hipStream_t)stream;
  const vt_tuning t = *tuning;
  // view records, kStageViews per staging launch (kernel arguments)
  uint64_t tiles = 0;
  vt::StagePack pack;
  for (uint32_t i0 = 0; i0 < n_views; i0 += vt::kStageViews) {
    pack.first = i0;
    pack.count = std::min(n_views - i0, vt::kStageViews);
    for (uint32_t k = 0; k < pack.count; k++)
      if ((st = view_record(views[i0 + k], t.tile, &tiles, &pack.v[k],
                            "vt_batch_render_noalloc")) != VT_OK)
        return st;
    hipLaunchKernelGGL(vt::k_stage_views, dim3(1), dim3(64), 0, s, sp->views, pack);
    VT_HIP(hipGetLastError());
  }
  bool steps_out = false;
  for (uint32_t i = 0; i < n_views; i++) steps_out |= views[i].frame.steps != nullptr;
  const hipError_t e = launch_batch(world, vt::shade_consts(p), sp->views, n_views,
                                    (uint32_t)tiles, sp->queues, t,
                                    (uint32_t)ctx->compute_units * t.blocks_per_cu, stats,
                                    steps_out, s);
  if (e != hipSuccess) return hip_fail(e, "vt_batch_render_noalloc: launch");
  return VT_OK;
}

}  // extern "C"
