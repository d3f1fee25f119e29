// vt_world.hip -- device-side construction of dense world layouts.
//
// The caller hands over a linear uint8 grid (x fastest).  On the device it is
// re-laid in Morton order (voxel (x,y,z) at interleave(x,y,z)), and the two
// occupancy levels are built from it: L0 = one uint64 per 64 consecutive
// codes (a 4^3 brick, bit = code & 63), L1 = one uint64 per 4096 codes (a
// 16^3 region, bit = brick index & 63 set when the brick has a solid voxel).
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

hipError_t launch_build_dense(const uint8_t* linear, uint32_t dim, uint8_t* vox, uint64_t* l0,
                              uint64_t* l1, hipStream_t stream) {
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
  return hipGetLastError();
}

}  // namespace vt
