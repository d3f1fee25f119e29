// vt_lambert.hip -- Lambert x visibility shading from stored hit records and
// its backward pass (the differentiable path; project-defined shading).
//
// A hit record (voxel id, t, face / shadow bit) fixes the discrete structure
// of a pixel, so shading and gradients need no traversal: the backward reads
// 12 B of record, 12 B of albedo and the upstream pixel, and adds 12 B of
// gradient.  Many pixels share a voxel: lanes hitting the same voxel in a row
// are summed in registers first (one float-atomic triple per run), and
// dL/dsun is reduced per wave before its three atomics.
#include <algorithm>
#include <cmath>

#include "vt_pixel.hpp"

namespace vt {

// In-kernel guard: a record whose voxel id lies outside the world reads as a
// miss, so no caller-supplied record can index past the world's arrays.
__device__ __forceinline__ bool record_hit(const DenseWorld& w, uint32_t voxel) {
  return voxel < w.dim * w.dim * w.dim;
}

// Hit record i in one 12-byte access (global_load_dwordx3 needs 4-byte
// alignment only): a wave reads its 64 records as one contiguous 768-byte run.
__device__ __forceinline__ uint3 load_record(const uint32_t* __restrict__ hits, uint32_t i) {
  return reinterpret_cast<const uint3*>(hits)[i];
}

__global__ __launch_bounds__(256) void k_lambert_shade(DenseWorld w, ShadeConsts sc,
                                                       const uint32_t* __restrict__ hits,
                                                       uint32_t npix, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix) return;
  const uint3 rec = load_record(hits, i);
  const uint32_t voxel = rec.x, info = rec.z;
  float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (record_hit(w, voxel)) {
    const int face = (int)(info & 7u);
    const bool occl = ((info >> 3) & 1u) != 0u;
    const LambertSun ls = lambert_sun(sc.sun);
    const float nd = lambert_nd(face, ls);
    const float fct = lambert_factor(sc.ambient, nd, occl);
    const uint32_t code = (sc.flags & VT_FLAG_ALBEDO_COMPACT) ? voxel_code(w, voxel) : 0u;
    const float* a = sc.albedo + 3 * (size_t)albedo_index(w, sc.flags, voxel, code);
    s = make_float4(a[0] * fct, a[1] * fct, a[2] * fct, 1.0f);
  }
  out[i] = s;
}

// Pixels per lane per block iteration of the backward: the loads of all
// of them (records, upstream, then the albedo gathers of the hits) are
// issued before any is used, so each lane keeps that many dependent gather
// chains in flight instead of one.
constexpr uint32_t kBwdUnroll = 4;

__global__ __launch_bounds__(256) void k_lambert_backward(DenseWorld w, ShadeConsts sc,
                                                          const uint32_t* __restrict__ hits,
                                                          uint32_t npix,
                                                          const float4* __restrict__ up,
                                                          uint32_t kind) {
  GradAcc acc = grad_acc();
  const LambertSun ls = lambert_sun(sc.sun);
  const bool compact = (sc.flags & VT_FLAG_ALBEDO_COMPACT) != 0u;
  const uint32_t span = 256u * kBwdUnroll;
  // grid-stride over block spans of 256 x kBwdUnroll pixels, the wave's
  // lanes on 64 consecutive pixels per sub-span (the bound is block-uniform,
  // so every wave reaches each flush_albedo); the launch caps the grid so
  // that the per-block loss / sun atomics stay few
  for (uint32_t base = blockIdx.x * span; base < npix; base += gridDim.x * span) {
    uint3 rec[kBwdUnroll];
    float4 u[kBwdUnroll];
#pragma unroll
    for (uint32_t j = 0; j < kBwdUnroll; j++) {
      const uint32_t i = base + j * 256u + threadIdx.x;
      rec[j] = make_uint3(0xFFFFFFFFu, 0u, 0u);
      u[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (i < npix) {
        rec[j] = load_record(hits, i);
        u[j] = up[i];
      }
    }
    uint32_t ai[kBwdUnroll];
    float A[kBwdUnroll][3];
#pragma unroll
    for (uint32_t j = 0; j < kBwdUnroll; j++) {
      ai[j] = kNoKey;
      A[j][0] = A[j][1] = A[j][2] = 0.0f;
      if (record_hit(w, rec[j].x)) {
        const uint32_t code = compact ? voxel_code(w, rec[j].x) : 0u;
        ai[j] = albedo_index(w, sc.flags, rec[j].x, code);
        const float* a = sc.albedo + 3 * (size_t)ai[j];
        A[j][0] = a[0];
        A[j][1] = a[1];
        A[j][2] = a[2];
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kBwdUnroll; j++) {
      if (ai[j] != kNoKey) {
        const uint32_t info = rec[j].z;
        const int face = (int)(info & 7u);
        const bool occl = ((info >> 3) & 1u) != 0u;
        const float nd = lambert_nd(face, ls);
        const float fct = lambert_factor(sc.ambient, nd, occl);
        float g[3];
        if (kind == VT_UPSTREAM_L2_TARGET) {
          const float r[3] = {A[j][0] * fct - u[j].x, A[j][1] * fct - u[j].y,
                              A[j][2] * fct - u[j].z};
          acc.loss += r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
          g[0] = 2.0f * r[0];
          g[1] = 2.0f * r[1];
          g[2] = 2.0f * r[2];
        } else {
          g[0] = u[j].x;
          g[1] = u[j].y;
          g[2] = u[j].z;
        }
        lambert_accumulate(ai[j], face, nd, occl, A[j], g, fct, ls, acc);
      } else if (kind == VT_UPSTREAM_L2_TARGET) {
        // sky pixel (or past the end: u = 0 adds nothing)
        acc.loss += u[j].x * u[j].x + u[j].y * u[j].y + u[j].z * u[j].z;
      }
      flush_albedo(sc, acc);
    }
  }
  flush_grad(sc, acc);
}

hipError_t launch_lambert_shade(const DenseWorld& w, const ShadeConsts& sc, const uint32_t* hits,
                                uint32_t npix, float4* out, hipStream_t stream) {
  hipLaunchKernelGGL(k_lambert_shade, dim3((npix + 255) / 256), dim3(256), 0, stream, w, sc, hits,
                     npix, out);
  return hipGetLastError();
}

hipError_t launch_lambert_backward(const DenseWorld& w, const ShadeConsts& sc,
                                   const uint32_t* hits, uint32_t npix, const float4* upstream,
                                   uint32_t kind, hipStream_t stream) {
  const uint32_t span = 256u * kBwdUnroll;
  const uint32_t blocks = std::min<uint32_t>((npix + span - 1) / span, 4096u);
  hipLaunchKernelGGL(k_lambert_backward, dim3(blocks), dim3(256), 0, stream, w, sc, hits, npix,
                     upstream, kind);
  return hipGetLastError();
}

}  // namespace vt

namespace {

using vt::aligned;
using vt::fail;
using vt::hip_fail;

int check_lambert_call(vt_ctx* ctx, const vt_world* world, const vt_hit_record* hits,
                       uint32_t npix, const vt_render_params* p, const char* who) {
  if (!ctx || !world || !hits || !p) return fail(VT_ERR_NULL_POINTER, "%s: null argument", who);
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "%s: world from another ctx", who);
  if (world->kind != WorldKind::Dense)
    return fail(VT_ERR_UNSUPPORTED, "%s: Lambert shading needs a dense world", who);
  if (npix == 0) return fail(VT_ERR_INVALID_ARGUMENT, "%s: npix is 0", who);
  if (!p->albedo) return fail(VT_ERR_NULL_POINTER, "%s: params->albedo required", who);
  if (!aligned(hits, 4) || !aligned(p->albedo, 4))
    return fail(VT_ERR_MISALIGNED, "%s: hits / albedo misaligned", who);
  if ((p->flags & VT_FLAG_ALBEDO_COMPACT) && !world->prefix)
    return fail(VT_ERR_UNSUPPORTED, "%s: build the compact index first", who);
  for (int i = 0; i < 3; i++)
    if (!std::isfinite(p->sun[i])) return fail(VT_ERR_INVALID_ARGUMENT, "%s: sun not finite", who);
  if (!vt::sun_normalizable(p->sun))
    return fail(VT_ERR_INVALID_ARGUMENT,
                "%s: Lambert shading normalises the sun: its squared length must be a nonzero "
                "finite binary32 value", who);
  return VT_OK;
}

}  // namespace

extern "C" {

int vt_world_build_compact_index(vt_ctx* ctx, vt_world* world, uint64_t* n_occupied) {
  if (!ctx || !world) return fail(VT_ERR_NULL_POINTER, "vt_world_build_compact_index: null");
  if (world->ctx != ctx) return fail(VT_ERR_WRONG_CONTEXT, "vt_world_build_compact_index");
  if (world->kind != WorldKind::Dense)
    return fail(VT_ERR_UNSUPPORTED, "vt_world_build_compact_index: dense worlds only");
  VT_HIP(hipSetDevice(ctx->device));
  if (!world->prefix) {
    const uint64_t nwords = (uint64_t)world->dim * world->dim * world->dim / 64u;
    const uint64_t ntiles = (nwords + 1023) / 1024;
    uint32_t *prefix = nullptr, *scratch = nullptr;
    hipError_t e = hipMalloc((void**)&prefix, nwords * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&scratch, (ntiles + 1) * 4);
    if (e == hipSuccess) e = vt::launch_compact_prefix(world->l0, nwords, prefix, scratch, nullptr);
    uint32_t total = 0;
    if (e == hipSuccess) e = hipMemcpy(&total, scratch + ntiles, 4, hipMemcpyDeviceToHost);
    vt::dev_free(scratch);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      vt::dev_free(prefix);
      return hip_fail(e, "vt_world_build_compact_index");
    }
    world->prefix = prefix;
    world->n_occupied = total;
    world->bytes += nwords * 4;
  }
  if (n_occupied) *n_occupied = world->n_occupied;
  return VT_OK;
}

int vt_lambert_shade(vt_ctx* ctx, const vt_world* world, const vt_hit_record* hits,
                     uint32_t npix, const vt_render_params* params, float* out, void* stream) {
  int st = check_lambert_call(ctx, world, hits, npix, params, "vt_lambert_shade");
  if (st != VT_OK) return st;
  if (!out) return fail(VT_ERR_NULL_POINTER, "vt_lambert_shade: out required");
  if (!aligned(out, 16)) return fail(VT_ERR_MISALIGNED, "vt_lambert_shade: out not 16B aligned");
  VT_HIP(hipSetDevice(ctx->device));
  hipError_t e = vt::launch_lambert_shade(vt::dense_view(world), vt::shade_consts(*params),
                                          reinterpret_cast<const uint32_t*>(hits), npix,
                                          reinterpret_cast<float4*>(out), (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "vt_lambert_shade: launch");
  return VT_OK;
}

int vt_lambert_backward(vt_ctx* ctx, const vt_world* world, const vt_hit_record* hits,
                        uint32_t npix, const vt_render_params* params, const float* upstream,
                        uint32_t upstream_kind, float* grad_albedo, float* grad_sun,
                        float* loss, void* stream) {
  int st = check_lambert_call(ctx, world, hits, npix, params, "vt_lambert_backward");
  if (st != VT_OK) return st;
  if (!upstream || !grad_albedo)
    return fail(VT_ERR_NULL_POINTER, "vt_lambert_backward: upstream and grad_albedo required");
  if (upstream_kind > VT_UPSTREAM_L2_TARGET)
    return fail(VT_ERR_INVALID_ARGUMENT, "vt_lambert_backward: unknown upstream kind %u",
                upstream_kind);
  if (!aligned(upstream, 16) || !aligned(grad_albedo, 4) || (grad_sun && !aligned(grad_sun, 4)) ||
      (loss && !aligned(loss, 4)))
    return fail(VT_ERR_MISALIGNED, "vt_lambert_backward: misaligned buffer");
  VT_HIP(hipSetDevice(ctx->device));
  vt::ShadeConsts sc = vt::shade_consts(*params);
  sc.grad_albedo = grad_albedo;
  sc.grad_sun = grad_sun;
  sc.loss = loss;
  hipError_t e = vt::launch_lambert_backward(
      vt::dense_view(world), sc, reinterpret_cast<const uint32_t*>(hits), npix,
      reinterpret_cast<const float4*>(upstream), upstream_kind, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "vt_lambert_backward: launch");
  return VT_OK;
}

}  // extern "C"
