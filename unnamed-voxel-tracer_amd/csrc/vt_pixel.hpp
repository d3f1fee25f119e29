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

// ---- Tile schedule (vt_frame.schedule) ----
// A frame's blocks are dispatched in launch order, round-robin over the 8
// XCDs (launch slot i runs on XCD i % 8), as wave slots free up.  In a
// terrain frame the blocks whose rays skim the terrain toward the horizon
// cost many times the others (a 1080p frame of the 512^3 world: 100-170 us
// per wave against 10-40 us elsewhere); in image order they start mid-launch
// and run on alone at its end (the launch's waves then hold 71% of the wave
// slots on average).  Launched most costly first, the cheap blocks fill in
// around them.  A block's cost is known once it has run: its first wave
// measures its time on the constant 100 MHz clock (the four waves of a block
// trace neighbouring tiles: at 1080p the first wave's time correlates 0.999
// with the block's longest), the block goes on its XCD's list of its log2
// cost class, and the next render into the same frame launches each XCD's
// blocks class by class from the costliest.  Each XCD keeps the image blocks
// the fixed order gives it (xcd_remap's strips: the L2 locality of that
// order; 1080p 243 us against 245 for one order over all XCDs), and only
// which of its launch slots renders which of them changes.
//
// Buffer (32-bit words; B = VT_SCHEDULE_BLOCKS(W, H) >= the blocks of either
// tile shape, C = ceil(B / 8) >= the blocks of one XCD): a head of
// kSchedHead words -- [0] the slot holding the last render's record, [1] the
// geometry tag of that render (bit 31 set, which no other word of the buffer
// reaches), [2] XCDs finished in the current render, [4 + 16x + k] the
// record's class counts, and from word 256 one counter of finished blocks
// per XCD, each in its own 128-byte line -- then three slots, each of 128
// class counters (XCD x, class k: one line each) and 128 lists of C block
// ids.  A render whose tag matches [1] launches XCD x's slots in the order
// of the record in slot r = [0] when x's counts sum to its block count (else
// in the fixed order: the same blocks either way), records into slot r+1
// (cleared by the render before) and clears slot r+2; any other tag (a
// zero-filled buffer, one last used at another size) means the fixed order,
// and the render clears all three slots.  A block reads the head in its
// first wave before a barrier; the first wave that finishes last copies the
// record's counts to the head (atomic reads: every block took its count
// before its XCD counter), writes [0] and [1] and resets the counters.
// Counters are spread over lines and counted per XCD because atomics on one
// address serialise: a counter taken by every wave of a 1080p launch (32,400)
// made the frame 480 us instead of 285.
constexpr uint32_t kSchedClasses = 16u;
constexpr uint32_t kSchedXcds = 8u;
constexpr uint32_t kSchedLists = kSchedXcds * kSchedClasses;  // 128
constexpr uint32_t kSchedLine = 32u;                          // words per 128-byte line
constexpr uint32_t kSchedCopy = 4u;                           // the record's counts in the head
constexpr uint32_t kSchedXcdDone = 256u;                      // per-XCD block counters
constexpr uint32_t kSchedHead = 512u;
constexpr uint32_t kSchedCounts = kSchedLine * kSchedLists;   // counters of a slot
static_assert(kSchedCopy + kSchedLists <= kSchedXcdDone &&
                  kSchedXcdDone + kSchedLine * kSchedXcds <= kSchedHead &&
                  kSchedCounts == 4096u,
              "VT_SCHEDULE_WORDS layout");

// The layout and tag of a launch's schedule, from the image size and tile
// shape (kernel arguments the render kernels have anyway) and the grid.
struct Sched {
  uint32_t* p;       // vt_frame.schedule, or null: the fixed order
  uint32_t C;        // list capacity: ceil(VT_SCHEDULE_BLOCKS(W, H) / 8)
  uint32_t tag;      // geometry tag of this launch (bit 31 set)
  uint32_t nblocks;  // blocks of this launch
};
__host__ __device__ __forceinline__ Sched sched_geometry(uint32_t* p, uint32_t W, uint32_t H,
                                                         uint32_t log2_tw, uint32_t nblocks) {
  Sched s;
  s.p = p;
  s.C = (VT_SCHEDULE_BLOCKS(W, H) + kSchedXcds - 1u) / kSchedXcds;
  s.tag = 0x80000000u | ((W * 0x9E3779B1u) ^ (H * 0x85EBCA77u) ^ (log2_tw * 0xC2B2AE3Du));
  s.nblocks = nblocks;
  return s;
}

// Launch slots on XCD x (blockIdx.x % 8 == x): the blocks the fixed order
// gives that XCD.
__host__ __device__ __forceinline__ uint32_t sched_xcd_blocks(uint32_t nblocks, uint32_t x) {
  return x < nblocks ? (nblocks - x + kSchedXcds - 1u) / kSchedXcds : 0u;
}

// How a block takes part: kSchedClear (the buffer is restarted) or
// kSchedRecord (its cost goes to slot rslot + 1).
enum : uint32_t { kSchedClear = 1, kSchedRecord = 2 };

__device__ __forceinline__ uint32_t* sched_slot(const Sched& s, uint32_t k) {
  return s.p + kSchedHead + k * (kSchedCounts + kSchedLists * s.C);
}
// The counter and the list of XCD x, class k in a slot.
__device__ __forceinline__ uint32_t* sched_count(uint32_t* slot, uint32_t x, uint32_t k) {
  return slot + (x * kSchedClasses + k) * kSchedLine;
}
__device__ __forceinline__ uint32_t* sched_list(uint32_t* slot, const Sched& s, uint32_t x,
                                                uint32_t k) {
  return slot + kSchedCounts + (x * kSchedClasses + k) * s.C;
}

// Block state kept in LDS from the block's start to the end of its last
// wave, so that the traversal in between holds no extra registers: the
// launch's schedule, the image block, the mode and read slot (mode | slot <<
// 2), the start time (low 32 bits of the 100 MHz clock) and the count of
// the block's waves that have finished.
struct SchedBlock {
  Sched s;
  uint32_t bid, mode_slot, t0, last, done;
};
// Waves of a render block (256 threads).
constexpr uint32_t kSchedWaves = 4;

// The image block this launch slot renders: rank blockIdx.x / 8 of its XCD's
// recorded order, or `fixed`.  The order is class by class from the
// costliest and, within a class, the order in which the blocks finished
// (1080p: 245 us per frame against 249 for the reverse).  The words read do
// not change during the launch, so every block of an XCD decides alike.
__device__ __forceinline__ uint32_t sched_block(const Sched& s, uint32_t fixed, uint32_t& mode,
                                                uint32_t& rslot) {
  rslot = 0u;
  const uint32_t rot = s.p[0], tag = s.p[1];
  if (tag != s.tag) {
    mode = kSchedClear;
    return fixed;
  }
  mode = kSchedRecord;
  rslot = rot % 3u;
  const uint32_t x = blockIdx.x % kSchedXcds;
  const uint32_t* cnt = s.p + kSchedCopy + x * kSchedClasses;
  uint32_t sum = 0u;
#pragma unroll
  for (uint32_t k = 0; k < kSchedClasses; ++k) sum += cnt[k];
  if (sum != sched_xcd_blocks(s.nblocks, x)) return fixed;
  uint32_t* R = sched_slot(s, rslot);
  uint32_t rank = blockIdx.x / kSchedXcds;
  for (int k = (int)kSchedClasses - 1; k >= 0; --k) {
    const uint32_t c = cnt[k];
    if (rank < c) {
      const uint32_t b = sched_list(R, s, x, (uint32_t)k)[rank];
      return b < s.nblocks ? b : fixed;
    }
    rank -= c;
  }
  return fixed;
}

// Start of a render kernel with a schedule (every thread): the image block
// this launch slot renders; the first wave decides and passes the block to
// the others through LDS.
__device__ __forceinline__ uint32_t sched_begin(const Sched& s, uint32_t fixed, uint32_t wave,
                                                SchedBlock& sb) {
  if (wave == 0u) {
    uint32_t mode, rslot;
    const uint32_t bid = sched_block(s, fixed, mode, rslot);
    sb.s = s;
    sb.bid = bid;
    sb.mode_slot = mode | (rslot << 2);
    sb.done = 0u;
    sb.t0 = (uint32_t)wall_clock64();
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(sb.bid);  // uniform: a scalar
}

// End of a render kernel with a schedule (every thread).  The block's last
// wave to finish does the bookkeeping, so that the block's cost is the time
// to its slowest wave (the first wave's own time misjudged blocks whose
// first wave ends early while another runs long, and such blocks launched
// last ran past the rest of the frame: 1080p 270 against 191 us).  Its lane
// 0: in record mode the block goes on its XCD's list of its log2 cost class and
// the launch's first blocks clear the next render's counters (in clear mode
// those of all three slots); then the block counts itself on its XCD's
// counter, and the block completing the last XCD has its wave publish the
// record.
__device__ __forceinline__ void sched_end(SchedBlock& sb, uint32_t lane) {
  uint32_t prev = 0u;
  if (lane == 0u)
    prev = __hip_atomic_fetch_add(&sb.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (__builtin_amdgcn_readfirstlane(prev) != kSchedWaves - 1u) return;
  const Sched s = sb.s;
  const uint32_t mode = sb.mode_slot & 3u, rslot = sb.mode_slot >> 2, bid = sb.bid;
  const uint32_t x = blockIdx.x % kSchedXcds;
  uint32_t* W = sched_slot(s, (rslot + 1u) % 3u);
  if (lane == 0u) {
    uint32_t last = 0u;
    if (mode == kSchedRecord) {
      // octaves of 100 MHz ticks from 2^5 (0.32 us) to 2^20 (10.5 ms), the
      // block's time from its start to the end of its last wave.  Measured
      // at 1080p with the first wave's time as the cost: octaves 243 us,
      // half octaves 243, half octaves of the first wave's most executed
      // steps instead of its time 248 (better at 1280x720: 138 against 148).
      const uint32_t d = (uint32_t)wall_clock64() - sb.t0;
      const uint32_t lg = 31u - (uint32_t)__clz(d | 1u);
      const uint32_t cls = lg < 5u ? 0u : min(lg - 5u, kSchedClasses - 1u);
      const uint32_t pos = atomicAdd(sched_count(W, x, cls), 1u);
      if (pos < s.C) sched_list(W, s, x, cls)[pos] = bid;
      uint32_t* Z = sched_slot(s, (rslot + 2u) % 3u);
      for (uint32_t q = blockIdx.x; q < kSchedLists; q += s.nblocks) Z[q * kSchedLine] = 0u;
    } else {
      for (uint32_t q = blockIdx.x; q < 3u * kSchedLists; q += s.nblocks)
        sched_slot(s, q / kSchedLists)[(q % kSchedLists) * kSchedLine] = 0u;
    }
    // the XCD counter is taken after the count returned (pos is used above)
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const uint32_t xcds = min(s.nblocks, kSchedXcds);
    if (atomicAdd(&s.p[kSchedXcdDone + x * kSchedLine], 1u) ==
            sched_xcd_blocks(s.nblocks, x) - 1u &&
        atomicAdd(&s.p[2], 1u) == xcds - 1u)
      last = 1u;
    sb.last = last;
  }
  __builtin_amdgcn_wave_barrier();
  if (sb.last) {  // the whole last wave of the last block
    for (uint32_t q = lane; q < kSchedLists; q += 64u)
      s.p[kSchedCopy + q] =
          mode == kSchedRecord
              ? __hip_atomic_load(&W[q * kSchedLine], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
              : 0u;
    if (lane < kSchedXcds) s.p[kSchedXcdDone + lane * kSchedLine] = 0u;
    if (lane == 0u) {
      s.p[0] = mode == kSchedRecord ? (rslot + 1u) % 3u : 2u;
      s.p[1] = s.tag;
      s.p[2] = 0u;
    }
  }
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
  const RayConsts pk = make_ray_consts(dx, dy, dz, w.mx, w.my, w.mz);
  if (!line_meets_reach(w, ex, ey, ez, pk)) {
    // this ray cannot reach a nonempty voxel: a miss, no steps (launches
    // whose step counts are observed pass a reach every line meets)
    h.hit = false;
    h.steps = 0;
  } else if (!dense_begin_k<kStats>(r, lk, w, ex, ey, ez, pk, sc.max_primary, h, cnt.l1f,
                                    cnt.l0f)) {
    r.blind = blind_steps(w, ex, ey, ez, pk);
    dense_run<kStats>(r, lk, w, h, cnt.l1f, cnt.l0f);
    // statistics of the steps executed (flat, free-region test): count
    // those taken without a lookup (VT_STAT_L1_FETCHES, unused by the flat
    // lookups otherwise)
    if constexpr (kStats && L::kPrefetch && L::kFree) cnt.l1f += (uint32_t)r.blind;
  }
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
