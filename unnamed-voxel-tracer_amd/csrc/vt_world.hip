// vt_world.hip -- device-side construction of dense world layouts.
//
// The caller hands over a linear uint8 grid (x fastest).  On the device it is
// re-laid in Morton order (voxel (x,y,z) at interleave(x,y,z)), and the
// occupancy levels are built from it: L0 = one uint64 per 64 consecutive
// codes (a 4^3 brick, bit = code & 63), L1 = one uint64 per 4096 codes (a
// 16^3 region, bit = brick index & 63 set when the brick has a solid voxel),
// L1b = one uint64 per 32768 codes (a 32^3 region, bit per 8^3 brick).
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
