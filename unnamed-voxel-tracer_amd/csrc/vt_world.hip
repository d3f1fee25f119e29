// vt_world.hip -- device-side construction of dense world layouts.
//
// The caller hands over a linear uint8 grid (x fastest).  On the device it is
// re-laid into 4^3 bricks of 64 contiguous bytes (one cache line holds a
// whole brick, so a ray stepping in any direction touches few lines), and the
// two occupancy levels are built: L0 = one uint64 per 4^3 brick (bit
// x + 4y + 16z), L1 = one uint64 per 16^3 region (bit per brick, same order).
#include "vt_internal.hpp"

namespace vt {

__global__ __launch_bounds__(256) void k_build_bricks(const uint8_t* __restrict__ lin, uint32_t dim,
                                                      uint8_t* __restrict__ vox,
                                                      uint64_t* __restrict__ l0) {
  const uint32_t nb = dim >> 2;
  const uint64_t nbricks = (uint64_t)nb * nb * nb;
  const uint64_t brick = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (brick >= nbricks) return;
  const uint32_t kx = (uint32_t)(brick % nb);
  const uint32_t ky = (uint32_t)((brick / nb) % nb);
  const uint32_t kz = (uint32_t)(brick / ((uint64_t)nb * nb));
  uint64_t mask = 0;
  uint32_t words[16];
#pragma unroll
  for (uint32_t z = 0; z < 4; z++) {
#pragma unroll
    for (uint32_t y = 0; y < 4; y++) {
      const uint64_t src = (uint64_t)kx * 4u + (uint64_t)dim * ((uint64_t)(ky * 4u + y) +
                                                                (uint64_t)dim * (kz * 4u + z));
      const uint32_t row = *reinterpret_cast<const uint32_t*>(lin + src);
      words[z * 4 + y] = row;
#pragma unroll
      for (uint32_t x = 0; x < 4; x++)
        if ((row >> (8 * x)) & 255u) mask |= 1ull << (x + 4 * y + 16 * z);
    }
  }
  uint4* dst = reinterpret_cast<uint4*>(vox + brick * 64u);
#pragma unroll
  for (int i = 0; i < 4; i++)
    dst[i] = make_uint4(words[4 * i + 0], words[4 * i + 1], words[4 * i + 2], words[4 * i + 3]);
  l0[brick] = mask;
}

__global__ __launch_bounds__(256) void k_build_regions(const uint64_t* __restrict__ l0,
                                                       uint32_t dim, uint64_t* __restrict__ l1) {
  const uint32_t nb = dim >> 2, nr = dim >> 4;
  const uint64_t nreg = (uint64_t)nr * nr * nr;
  const uint64_t reg = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (reg >= nreg) return;
  const uint32_t rx = (uint32_t)(reg % nr);
  const uint32_t ry = (uint32_t)((reg / nr) % nr);
  const uint32_t rz = (uint32_t)(reg / ((uint64_t)nr * nr));
  uint64_t mask = 0;
  for (uint32_t z = 0; z < 4; z++)
    for (uint32_t y = 0; y < 4; y++)
      for (uint32_t x = 0; x < 4; x++) {
        const uint64_t b = ((uint64_t)(rz * 4 + z) * nb + (ry * 4 + y)) * nb + (rx * 4 + x);
        if (l0[b]) mask |= 1ull << (x + 4 * y + 16 * z);
      }
  l1[reg] = mask;
}

hipError_t launch_build_dense(const uint8_t* linear, uint32_t dim, uint8_t* vox, uint64_t* l0,
                              uint64_t* l1, hipStream_t stream) {
  const uint64_t nb = dim >> 2, nr = dim >> 4;
  const uint64_t nbricks = nb * nb * nb, nreg = nr * nr * nr;
  hipLaunchKernelGGL(k_build_bricks, dim3((uint32_t)((nbricks + 255) / 256)), dim3(256), 0, stream,
                     linear, dim, vox, l0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_build_regions, dim3((uint32_t)((nreg + 255) / 256)), dim3(256), 0, stream,
                     l0, dim, l1);
  return hipGetLastError();
}

}  // namespace vt
