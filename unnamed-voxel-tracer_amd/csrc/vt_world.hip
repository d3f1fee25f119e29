// vt_world.hip -- device-side construction of dense world layouts.
//
// The caller hands over a linear uint8 grid (x fastest).  On the device it is
// re-laid in Morton order (voxel (x,y,z) at interleave(x,y,z)), and the
// occupancy levels are built from it: L0 = one uint64 per 64 consecutive
// codes (a 4^3 brick, bit = code & 63), L1 = one uint64 per 4096 codes (a
// 16^3 region, bit = brick index & 63 set when the brick has a solid voxel),
// L1b = one uint64 per 32768 codes (a 32^3 region, bit per 8^3 brick).
#include <algorithm>
#include <vector>

#include "vt_internal.hpp"

namespace vt {

// One thread per 4 consecutive codes (one 32-bit store).
__global__ __launch_bounds__(256) void k_build_morton(const uint8_t* __restrict__ lin,
                                                      uint32_t dim, uint8_t* __restrict__ vox) {
  const uint64_t n = (uint64_t)dim * dim * dim;
  const uint64_t c0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4u;
  if (c0 >= n) return;
  uint32_t word = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; k++) {
    const uint32_t c = (uint32_t)c0 + k;
    const uint32_t x = compact1by2(c), y = compact1by2(c >> 1), z = compact1by2(c >> 2);
    const uint32_t b = lin[(uint64_t)x + (uint64_t)dim * ((uint64_t)y + (uint64_t)dim * z)];
    word |= b << (8 * k);
  }
  *reinterpret_cast<uint32_t*>(vox + c0) = word;
}

// One thread per L0 word: 64 consecutive Morton bytes.
__global__ __launch_bounds__(256) void k_build_l0(const uint8_t* __restrict__ vox, uint64_t nwords,
                                                  uint64_t* __restrict__ l0) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  const uint4* src = reinterpret_cast<const uint4*>(vox + i * 64u);
  uint64_t mask = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint4 v = src[q];
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        if ((w[j] >> (8 * b)) & 255u) mask |= 1ull << (q * 16 + j * 4 + b);
  }
  l0[i] = mask;
}

// One thread per L1 word: 64 consecutive L0 words.
__global__ __launch_bounds__(256) void k_build_l1(const uint64_t* __restrict__ l0, uint64_t nwords,
                                                  uint64_t* __restrict__ l1) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  uint64_t mask = 0;
  for (uint32_t j = 0; j < 64; j++)
    if (l0[i * 64u + j]) mask |= 1ull << j;
  l1[i] = mask;
}

// The 8^3-brick level: one uint64 per 32^3 region (32768 codes), bit j set
// when 8^3 brick 64 i + j (L0 words 8 (64 i + j) .. + 7) holds a solid
// voxel.  A 16^3 world has one partial word (8 bricks).  One wave per word.
__global__ __launch_bounds__(256) void k_build_l1b(const uint64_t* __restrict__ l0, uint64_t n0,
                                                   const uint32_t* __restrict__ words,
                                                   uint32_t nwords, uint64_t* __restrict__ l1b) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  if (wave >= nwords) return;  // whole waves leave together
  const uint32_t word = words ? words[wave] : wave;
  const uint64_t first = ((uint64_t)word * 64u + lane) * 8u;
  uint64_t any = 0;
  if (first < n0)
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) any |= l0[first + k];
  const uint64_t m = __ballot(any != 0ull);
  if (lane == 0) l1b[word] = m;
}

hipError_t launch_build_l1b(const uint64_t* l0, uint64_t n0, const uint32_t* words,
                            uint32_t nwords, uint64_t* l1b, hipStream_t stream) {
  if (nwords == 0) return hipSuccess;
  hipLaunchKernelGGL(k_build_l1b, dim3((nwords + 3u) / 4u), dim3(256), 0, stream, l0, n0, words,
                     nwords, l1b);
  return hipGetLastError();
}

// ---- compact albedo index: exclusive prefix sum of popcount(L0 word) ----
// Three passes: per-tile sums (1024 words per 256-thread block), one block
// scanning the tile sums, then each tile scanning its own words.
constexpr uint32_t kScanTile = 1024;

__device__ __forceinline__ uint32_t block_exclusive_scan_256(uint32_t v, uint32_t* sh,
                                                             uint32_t& total) {
  const uint32_t t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t a = t >= off ? sh[t - off] : 0u;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  total = sh[255];
  const uint32_t incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_popc_tile_sums(const uint64_t* __restrict__ l0,
                                                        uint64_t nwords,
                                                        uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[256];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * 4u;
  uint32_t s = 0;
  for (uint32_t k = 0; k < 4; k++)
    if (base + k < nwords) s += (uint32_t)__popcll(l0[base + k]);
  uint32_t total;
  (void)block_exclusive_scan_256(s, sh, total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_tile_sums(uint32_t* __restrict__ sums,
                                                         uint32_t ntiles,
                                                         uint32_t* __restrict__ total) {
  __shared__ uint32_t buf[1024];
  const uint32_t t = threadIdx.x;
  uint32_t carry = 0;
  for (uint32_t c = 0; c < ntiles; c += 1024u) {
    const uint32_t i = c + t;
    const uint32_t v = i < ntiles ? sums[i] : 0u;
    buf[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
      const uint32_t a = t >= off ? buf[t - off] : 0u;
      __syncthreads();
      buf[t] += a;
      __syncthreads();
    }
    if (i < ntiles) sums[i] = carry + buf[t] - v;
    carry += buf[1023];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

__global__ __launch_bounds__(256) void k_prefix_words(const uint64_t* __restrict__ l0,
                                                      uint64_t nwords,
                                                      const uint32_t* __restrict__ tile_off,
                                                      uint32_t* __restrict__ prefix) {
  __shared__ uint32_t sh[256];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + threadIdx.x * 4u;
  uint32_t pc[4], s = 0;
  for (uint32_t k = 0; k < 4; k++) {
    pc[k] = base + k < nwords ? (uint32_t)__popcll(l0[base + k]) : 0u;
    s += pc[k];
  }
  uint32_t total;
  uint32_t run = tile_off[blockIdx.x] + block_exclusive_scan_256(s, sh, total);
  for (uint32_t k = 0; k < 4; k++) {
    if (base + k < nwords) prefix[base + k] = run;
    run += pc[k];
  }
}

hipError_t launch_compact_prefix(const uint64_t* l0, uint64_t nwords, uint32_t* prefix,
                                 uint32_t* scratch, hipStream_t stream) {
  const uint32_t ntiles = (uint32_t)((nwords + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_popc_tile_sums, dim3(ntiles), dim3(256), 0, stream, l0, nwords, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, stream, scratch, ntiles,
                     scratch + ntiles);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_prefix_words, dim3(ntiles), dim3(256), 0, stream, l0, nwords, scratch,
                     prefix);
  return hipGetLastError();
}

hipError_t launch_build_dense(const uint8_t* linear, uint32_t dim, uint8_t* vox, uint64_t* l0,
                              uint64_t* l1, uint64_t* l1b, hipStream_t stream) {
  const uint64_t n = (uint64_t)dim * dim * dim;  // a power of two >= 4096
  const uint64_t n0 = n / 64u, n1 = n / 4096u;
  hipLaunchKernelGGL(k_build_morton, dim3((uint32_t)((n / 4u + 255) / 256)), dim3(256), 0, stream,
                     linear, dim, vox);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_build_l0, dim3((uint32_t)((n0 + 255) / 256)), dim3(256), 0, stream, vox, n0,
                     l0);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_build_l1, dim3((uint32_t)((n1 + 255) / 256)), dim3(256), 0, stream, l0, n1,
                     l1);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_build_l1b(l0, n0, nullptr, (uint32_t)l1b_words(dim), l1b, stream);
}

// ---- free-region map: where a ray's remaining path is provably empty ----
// Regions are 16^3 blocks (one L1 word; R = max(dim / 16, 1) per axis).  For
// each of the 8 direction octants o, free_o(r) = no region r' at or ahead of
// r on every axis in o's travel direction (the forward orthant, r itself
// included) holds a nonempty block.  Bit r (region Morton code) of octant
// o's bitmap is free_o(r).  The traversal loops read it to end a ray that can
// hit nothing more.

// One thread per region (linear index rx + R (ry + R rz)); writes whether the
// region holds a nonempty brick into all 8 octant planes of tmp.
__global__ __launch_bounds__(256) void k_free_regions(const uint64_t* __restrict__ l1, uint32_t R,
                                                      uint8_t* __restrict__ tmp) {
  const uint32_t n = R * R * R;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t rx = i % R, ry = (i / R) % R, rz = i / (R * R);
  const uint8_t v = l1[morton3(rx, ry, rz)] != 0ull ? 1 : 0;
  for (uint32_t o = 0; o < 8; o++) tmp[(size_t)o * n + i] = v;
}

// One thread per (octant, line along `axis`): each region becomes the OR of
// itself and every region ahead of it on that axis in the octant's travel
// direction.  Three passes (x, y, z) give the forward-orthant OR.
__global__ __launch_bounds__(256) void k_free_scan(uint8_t* __restrict__ tmp, uint32_t R,
                                                   int axis) {
  const uint32_t lines = R * R, n = lines * R;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8u * lines) return;
  const uint32_t o = t / lines, l = t % lines;
  const uint32_t u = l % R, v = l / R;
  size_t base, stride;
  if (axis == 0) {
    base = (size_t)R * (u + (size_t)R * v);
    stride = 1;
  } else if (axis == 1) {
    base = u + (size_t)lines * v;
    stride = R;
  } else {
    base = u + (size_t)R * v;
    stride = lines;
  }
  uint8_t* p = tmp + (size_t)o * n + base;
  uint8_t acc = 0;
  if (((o >> axis) & 1u) == 0u) {  // positive: ahead = higher index
    for (int k = (int)R - 1; k >= 0; --k) {
      acc |= p[(size_t)k * stride];
      p[(size_t)k * stride] = acc;
    }
  } else {
    for (uint32_t k = 0; k < R; ++k) {
      acc |= p[(size_t)k * stride];
      p[(size_t)k * stride] = acc;
    }
  }
}

// One thread per (octant, bitmap word): bit j of word w = free_o(region with
// Morton code 32 w + j).
__global__ __launch_bounds__(256) void k_free_pack(const uint8_t* __restrict__ tmp, uint32_t R,
                                                   uint32_t words, uint32_t* __restrict__ fm) {
  const uint32_t n = R * R * R;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8u * words) return;
  const uint32_t o = t / words, wi = t % words;
  uint32_t bits = 0;
  for (uint32_t j = 0; j < 32; j++) {
    const uint32_t rc = wi * 32u + j;
    if (rc >= n) break;
    const uint32_t x = compact1by2(rc), y = compact1by2(rc >> 1), z = compact1by2(rc >> 2);
    if (!tmp[(size_t)o * n + x + (size_t)R * (y + (size_t)R * z)]) bits |= 1u << j;
  }
  fm[t] = bits;
}

uint32_t free_regions(uint32_t dim) { return dim >= 16u ? dim / 16u : 1u; }

uint32_t free_words_per_octant(uint32_t dim) {
  const uint32_t R = free_regions(dim);
  return (R * R * R + 31u) / 32u;
}

hipError_t launch_build_freemap(const uint64_t* l1, uint32_t dim, uint8_t* tmp, uint32_t* fm,
                                hipStream_t stream) {
  const uint32_t R = free_regions(dim), n = R * R * R, words = free_words_per_octant(dim);
  hipLaunchKernelGGL(k_free_regions, dim3((n + 255u) / 256u), dim3(256), 0, stream, l1, R, tmp);
  hipError_t e = hipGetLastError();
  for (int axis = 0; axis < 3 && e == hipSuccess; axis++) {
    hipLaunchKernelGGL(k_free_scan, dim3((8u * R * R + 255u) / 256u), dim3(256), 0, stream, tmp,
                       R, axis);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_free_pack, dim3((8u * words + 255u) / 256u), dim3(256), 0, stream, tmp, R,
                     words, fm);
  return hipGetLastError();
}

// Brick bits of a Morton-ordered byte array (the brickmap occupancy bytes):
// one uint64 per 4096 codes, bit j set when bytes [64 (64 w + j), +64) hold
// a nonzero byte.  One wave per word; an 8^3 map has one partial word.
__global__ __launch_bounds__(256) void k_brick_bits(const uint8_t* __restrict__ bytes,
                                                    uint64_t nbytes, uint32_t nwords,
                                                    uint64_t* __restrict__ l1) {
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63u;
  if (wave >= nwords) return;  // whole waves leave together
  const uint64_t first = ((uint64_t)wave * 64u + lane) * 64u;
  bool any = false;
  if (first < nbytes) {
    const uint4* s = reinterpret_cast<const uint4*>(bytes + first);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 v = s[q];
      any |= (v.x | v.y | v.z | v.w) != 0u;
    }
  }
  const uint64_t m = __ballot(any);
  if (lane == 0) l1[wave] = m;
}

uint32_t brick_bit_words(uint32_t dim) {
  const uint64_t n = (uint64_t)dim * dim * dim;
  return n >= 4096u ? (uint32_t)(n / 4096u) : 1u;
}

hipError_t launch_brick_bits(const uint8_t* bytes, uint32_t dim, uint64_t* l1,
                             hipStream_t stream) {
  const uint32_t nwords = brick_bit_words(dim);
  hipLaunchKernelGGL(k_brick_bits, dim3((nwords + 3u) / 4u), dim3(256), 0, stream, bytes,
                     (uint64_t)dim * dim * dim, nwords, l1);
  return hipGetLastError();
}

hipError_t rebuild_freemap(vt_world* w, hipStream_t stream) {
  w->free_valid = false;
  if (!w->freemap) return hipSuccess;
  if (w->kind == WorldKind::Dense) return launch_build_freemap(w->l1, w->dim, w->free_tmp,
                                                               w->freemap, stream);
  hipError_t e = launch_brick_bits(w->occ, w->dim, w->brick_bits, stream);
  if (e != hipSuccess) return e;
  return launch_build_freemap(w->brick_bits, w->dim, w->free_tmp, w->freemap, stream);
}

hipError_t finish_freemap(vt_world* w) {
  w->free_valid = false;
  w->free_frac = 0.0f;
  if (!w->freemap) return hipSuccess;
  const uint32_t words = free_words_per_octant(w->dim), R = free_regions(w->dim);
  std::vector<uint32_t> h(8u * (size_t)words);
  const hipError_t e = hipMemcpy(h.data(), w->freemap, h.size() * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  for (uint32_t o = 0; o < 8; o++) {
    uint64_t n = 0;
    for (uint32_t i = 0; i < words; i++) n += (uint64_t)__builtin_popcount(h[o * words + i]);
    w->free_frac = std::max(w->free_frac, (float)((double)n / ((double)R * R * R)));
  }
  w->free_valid = true;
  return hipSuccess;
}

bool use_free_map(const vt_world* w, const ShadeConsts& sc) {
  if (!w->freemap || !w->free_valid || w->free_frac <= 0.0f) return false;
  if (sc.flags & VT_FLAG_FREE_MAP_OFF) return false;
  if (sc.flags & VT_FLAG_FREE_MAP_ON) return true;
  return (sc.max_primary >= VT_FREE_MIN_CAP || sc.max_shadow >= VT_FREE_MIN_CAP) &&
         w->free_frac * VT_FREE_MIN_SHARE >= 1.0f;
}

// One thread per block of the layers y < y_end of a brickmap world (linear
// order x + dim*(y + y_end*z)).
__global__ __launch_bounds__(256) void k_build_occupancy(BrickmapWorld w, uint8_t* __restrict__ occ,
                                                         uint32_t y_end) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = (uint64_t)w.dim * y_end * w.dim;
  if (i >= n) return;
  const uint32_t x = (uint32_t)(i % w.dim), y = (uint32_t)((i / w.dim) % y_end),
                 z = (uint32_t)(i / ((uint64_t)w.dim * y_end));
  occ[morton3(x, y, z)] = bm_block(w, (int)x, (int)y, (int)z) != 0u ? 1 : 0;
}

hipError_t launch_build_occupancy(const BrickmapWorld& w, uint8_t* occ, hipStream_t stream,
                                  uint32_t y_end) {
  if (y_end == 0 || y_end > w.dim) y_end = w.dim;
  const uint64_t n = (uint64_t)w.dim * y_end * w.dim;
  hipLaunchKernelGGL(k_build_occupancy, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream, w,
                     occ, y_end);
  return hipGetLastError();
}

}  // namespace vt
