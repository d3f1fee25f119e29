// vt_internal.hpp -- host-side objects behind the C-ABI handles and the
// launchers each kernel translation unit exports.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/vt.h"
#include "vt_trace.hpp"

// Device staging owned by a context, allocated once by vt_create: the pick
// result and the coordinates / values of vt_world_get_voxels (in chunks of
// kCtxStageVoxels), so neither allocates per call.  Calls that use it
// return only after their stream work is done, so calls made one after
// another on the context's thread never overlap on it.
constexpr size_t kCtxStageBytes = 256u * 1024u;
constexpr uint32_t kCtxStageVoxels = (uint32_t)(kCtxStageBytes / 16u);  // 12 B xyz + 4 B value

struct vt_ctx {
  int device;
  int compute_units;
  char name[256];
  void* stage;  // kCtxStageBytes of device memory, 256-byte aligned
};

enum class WorldKind { Dense, Brickmap };

struct vt_world {
  vt_ctx* ctx;
  WorldKind kind;
  uint32_t dim;
  size_t bytes;
  // dense (Morton order)
  uint8_t* vox;
  uint64_t* l0;
  uint64_t* l1;
  uint64_t* l1b;  // 8^3-brick level, l1b_words(dim) words
  uint32_t* palette;
  // brickmap
  uint32_t* chunks;
  uint32_t* pool;
  uint32_t* atlas;
  size_t pool_words;     // words of allocated blocks (512 per block)
  size_t pool_capacity;  // words the device pool holds
  uint32_t* chunk_mirror;  // host copy of the chunk table (edit allocation)
  uint8_t* occ;            // block occupancy in Morton order (power-of-two dim)
  uint64_t* brick_bits;    // brickmap: bit per 4^3 brick of blocks, from occ
  // free-region map (dense worlds and power-of-two brickmap worlds)
  uint32_t* freemap;       // 8 octants x free_words_per_octant(dim) words
  uint8_t* free_tmp;       // 8 x free_regions(dim)^3 bytes of build scratch
  bool free_valid;         // the map matches the occupancy (finish_freemap)
  float free_frac;         // largest share of free regions over the 8 octants
  // the box of the nonempty voxels (dense) or blocks (reference layout),
  // inclusive; grown by edits, never shrunk; box_any = false while the
  // world has none
  int32_t box_lo[3], box_hi[3];
  bool box_any;
  // compact albedo index of a dense world (built on request)
  uint32_t* prefix;     // occupied voxels before each L0 word
  uint64_t n_occupied;
};

namespace vt {

// Error reporting shared by the C-ABI translation units: record a detail
// message for vt_last_error() and return the status.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int hip_fail(hipError_t e, const char* what);
bool aligned(const void* p, size_t a);
void dev_free(void* p);

#define VT_HIP(call)                                       \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return ::vt::hip_fail(e_, #call); \
  } while (0)

// Per-launch shading constants.
struct ShadeConsts {
  float sun[3];
  RayConsts shadow;  // the shadow rays' direction constants (set by the dense launchers)
  int max_primary, max_shadow;
  uint32_t flags;
  float ambient;
  const float* albedo;
  float* grad_albedo;  // VT_FLAG_GRAD_L2 accumulators
  float* grad_sun;
  float* loss;
};

// Output pointers (any may be null).
struct alignas(8) FrameOut {
  float4* shaded;
  uint32_t* hits;  // 3 words per pixel
  float4* color;
  float4* normal;
  float4* position;
  float4* illum;
  uint32_t* steps;
  unsigned long long* stats;
  const float4* target;  // VT_FLAG_GRAD_L2 input
  uint32_t* display;     // VT_FLAG_PRESENT RGBA8 output
};

// Wave tile: 64 lanes cover tile_w x (64 / tile_w) pixels; a 256-thread
// block covers 2 x 2 wave tiles.
struct TileShape {
  uint32_t log2_w;  // 3 (8x8), 4 (16x4) or 5 (32x2)
};

DenseWorld dense_view(const vt_world* w);
void scan_dense_box(const uint8_t* grid, uint32_t dim, vt_world* w);
void scan_brickmap_box(const uint32_t* chunks, const uint32_t* pool, uint32_t dim, vt_world* w);
void grow_box(vt_world* w, int32_t x, int32_t y, int32_t z);
void reach_of(const vt_world* w, float* lo, float* hi);
BrickmapWorld brickmap_view(const vt_world* w);

// Takes the device address of the no-allocation scratch array on `device`
// (the current device) once per device; called by vt_create.
int noalloc_init(int device);

// Host-side validation and translation of the C-ABI structs.
int check_image_size(uint32_t W, uint32_t H, const char* who);
int check_params(const vt_render_params& p, const vt_world* world, const char* who);
// The Lambert modes divide by |sun|: false when its binary32 squared length
// is 0 (or overflows), where that division has no finite result.
bool sun_normalizable(const float sun[3]);
int check_frame(const vt_frame* frame, uint32_t flags, const char* who);
ViewConsts view_consts(const vt_camera* camera, uint32_t W, uint32_t H);
ShadeConsts shade_consts(const vt_render_params& p);
FrameOut frame_out(const vt_frame* frame);

// `schedule`: the frame's tile schedule buffer (vt_frame.schedule) or null.
hipError_t launch_render_dense(const DenseWorld& w, uint32_t traversal, const ViewConsts& v,
                               const ShadeConsts& sc, const FrameOut& out, uint32_t* schedule,
                               TileShape tile, uint32_t compute_units, hipStream_t stream);
hipError_t launch_render_brickmap(const BrickmapWorld& w, const ViewConsts& v,
                                  const ShadeConsts& sc, const FrameOut& out, uint32_t* schedule,
                                  TileShape tile, uint32_t compute_units, hipStream_t stream);

// World construction kernels.
// occ[morton(b)] = (map_getVoxel(b) != 0) for every block b of a brickmap
// world with y < y_end (0: all layers).
hipError_t launch_build_occupancy(const BrickmapWorld& w, uint8_t* occ, hipStream_t stream,
                                  uint32_t y_end = 0);
// Grows a brickmap world's block pool to at least `blocks` blocks by
// doubling (GpuBlockAllocator.alloc, gpu_block_allocator.zig:20-30), keeping
// the contents and zeroing the new blocks; `who` names the caller in errors.
int grow_pool(vt_world* w, uint64_t blocks, hipStream_t s, const char* who);
hipError_t launch_build_dense(const uint8_t* linear, uint32_t dim, uint8_t* vox, uint64_t* l0,
                              uint64_t* l1, uint64_t* l1b, hipStream_t stream);
// Words of the 8^3-brick level of a dim^3 world: one per 32^3 region (one
// partial word for dim 16).
inline uint64_t l1b_words(uint32_t dim) {
  const uint64_t n = (uint64_t)dim * dim * dim / 32768u;
  return n ? n : 1u;
}
// Rebuilds the listed L1b words (all l1b_words(dim) when words is null)
// from the L0 words, n0 of them.
hipError_t launch_build_l1b(const uint64_t* l0, uint64_t n0, const uint32_t* words,
                            uint32_t nwords, uint64_t* l1b, hipStream_t stream);
// Free-region map: regions of 16^3 blocks, R = free_regions(dim) per axis;
// bit r of octant o's bitmap (free_words_per_octant(dim) words) is set when
// no region at or ahead of r on every axis in o's travel direction holds a
// nonempty block.  Built from the brick bits (a dense world's L1 words, a
// brickmap world's brick_bits) with 8 R^3 bytes of scratch.
uint32_t free_regions(uint32_t dim);
uint32_t free_words_per_octant(uint32_t dim);
hipError_t launch_build_freemap(const uint64_t* l1, uint32_t dim, uint8_t* tmp, uint32_t* fm,
                                hipStream_t stream);
// Rebuilds a world's free-region map from its current occupancy (a dense
// world's L1 words; a brickmap world's occupancy bytes); no-op without one.
// Clears free_valid: the caller calls finish_freemap once the stream has
// synchronised without error, and until then the renderers do not read it.
hipError_t rebuild_freemap(vt_world* w, hipStream_t stream);
// After a rebuild completed: reads the map back to learn the largest share
// of free regions over the octants (free_frac) and marks it valid.
hipError_t finish_freemap(vt_world* w);
// Whether a launch with these constants tests the free-region map: the map
// is valid and has a free region, and VT_FLAG_FREE_MAP_ON, or neither
// FREE_MAP flag, a step cap of at least VT_FREE_MIN_CAP and at least
// 1/VT_FREE_MIN_SHARE of the regions free in some octant.  (Frames that
// store step counts or statistics are excluded by the launchers.)
bool use_free_map(const vt_world* w, const ShadeConsts& sc);
// Whether vt_render makes the test for this world, parameters and frame:
// use_free_map, and no stored step counts; on dense worlds the flat
// traversal, and statistics only with VT_FLAG_FREE_MAP_ON (they then count
// the steps executed); on the reference layout occupancy bytes and no
// statistics.
bool render_free_map(const vt_world* w, const vt_render_params& p, const vt_frame& f);
// Brick bits (one uint64 per 4096 codes) of a Morton-ordered byte array of
// dim^3 bytes; brick_bit_words(dim) words.
uint32_t brick_bit_words(uint32_t dim);
hipError_t launch_brick_bits(const uint8_t* bytes, uint32_t dim, uint64_t* l1,
                             hipStream_t stream);
// Exclusive prefix sum of popcount(l0[i]) into prefix[0..nwords); scratch
// holds ceil(nwords/1024) + 1 words; *total receives the sum (device).
hipError_t launch_compact_prefix(const uint64_t* l0, uint64_t nwords, uint32_t* prefix,
                                 uint32_t* scratch, hipStream_t stream);

// Lambert shading / backward from stored hit records.
hipError_t launch_present(const float4* composite, uint32_t W, uint32_t H, float4* out,
                          uint32_t* display, hipStream_t stream);
hipError_t launch_lambert_shade(const DenseWorld& w, const ShadeConsts& sc, const uint32_t* hits,
                                uint32_t npix, float4* out, hipStream_t stream);
hipError_t launch_lambert_backward(const DenseWorld& w, const ShadeConsts& sc,
                                   const uint32_t* hits, uint32_t npix, const float4* upstream,
                                   uint32_t kind, hipStream_t stream);

}  // namespace vt
