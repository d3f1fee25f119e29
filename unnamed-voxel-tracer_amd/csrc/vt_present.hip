// vt_present.hip -- the present pass over a composite image
// (blit.fragment.glsl:32-35: crosshair, vignette, store to an rgba8 target).
//
// A streaming kernel: 16 B read and 16 B (RGBA32F) and/or 4 B (RGBA8) written
// per pixel, one pixel per lane with 16-byte vector accesses, so consecutive
// lanes touch consecutive 16-byte texels -- HBM bound.  vt_render with
// VT_FLAG_PRESENT applies the same present_pixel() inside the render kernel
// and saves this pass.
#include "vt_internal.hpp"

namespace vt {

__global__ __launch_bounds__(256) void k_present(const float4* __restrict__ composite,
                                                 uint32_t W, uint32_t H, float4* __restrict__ out,
                                                 uint32_t* __restrict__ display) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)W * H) return;
  const uint32_t y = (uint32_t)(i / W), x = (uint32_t)(i - (uint64_t)y * W);
  const float4 c = present_pixel(composite[i], x, y, W, H);
  if (out) out[i] = c;
  if (display) display[i] = pack_unorm8(c);
}

hipError_t launch_present(const float4* composite, uint32_t W, uint32_t H, float4* out,
                          uint32_t* display, hipStream_t stream) {
  const uint64_t n = (uint64_t)W * H;
  hipLaunchKernelGGL(k_present, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream,
                     composite, W, H, out, display);
  return hipGetLastError();
}

}  // namespace vt

extern "C" int vt_present(vt_ctx* ctx, const float* composite, uint32_t width, uint32_t height,
                          float* out, uint32_t* display, void* stream) {
  using vt::aligned;
  using vt::fail;
  if (!ctx || !composite) return fail(VT_ERR_NULL_POINTER, "vt_present: null argument");
  if (!out && !display) return fail(VT_ERR_NULL_POINTER, "vt_present: no output");
  int st = vt::check_image_size(width, height, "vt_present");
  if (st != VT_OK) return st;
  if (!aligned(composite, 16) || (out && !aligned(out, 16)) || (display && !aligned(display, 4)))
    return fail(VT_ERR_MISALIGNED, "vt_present: misaligned image");
  VT_HIP(hipSetDevice(ctx->device));
  hipError_t e = vt::launch_present(reinterpret_cast<const float4*>(composite), width, height,
                                    reinterpret_cast<float4*>(out), display, (hipStream_t)stream);
  if (e != hipSuccess) return vt::hip_fail(e, "vt_present: launch");
  return VT_OK;
}
