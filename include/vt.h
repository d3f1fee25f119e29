/*
 * vt.h -- C-ABI of the MI355X voxel ray-traversal library (libvt.so).
 *
 * This is the drop-in boundary for the reference engine's render dispatch:
 * what `Game.render` (src/game.zig:232-256) binds and dispatches through
 * `ComputePipeline.dispatch` (src/engine/graphics/shader.zig:113-117) --
 * camera UBO 8, block pool SSBO 9, chunk table SSBO 10, model atlas image 6,
 * G-buffer images 0-3 -- becomes a world handle, a camera struct and caller
 * owned device output buffers, traced by hand-written gfx950 HIP kernels.
 *
 * Conventions
 *  - Every entry point returns an int status (VT_OK == 0); nothing throws or
 *    aborts across the ABI.  vt_status_string() names a status.
 *  - Host arrays handed to vt_world_create_* are copied; the library owns all
 *    world device memory until vt_world_destroy().
 *  - Output buffers in vt_frame are caller-owned DEVICE pointers.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Work is enqueued asynchronously and in order on that stream, like the
 *    reference's in-order GL queue plus glMemoryBarrier (shader.zig:116).
 *  - One vt_ctx per host thread; a vt_ctx is bound to one HIP device.
 *
 * Reference semantics ("REF"): the traversal restates traceMap
 * (assets/shaders/map.glsl:83-168) operation by operation in IEEE-754 binary32
 * with round-to-nearest-even, no fused multiply-add, correctly rounded
 * division and square root, so hits are bit-exact against the CPU
 * restatement.
 */
#ifndef VT_H
#define VT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  VT_OK = 0,
  VT_ERR_INVALID_ARGUMENT = 1, /* a size, flag or enum value is out of range   */
  VT_ERR_NULL_POINTER = 2,     /* a required pointer is NULL                   */
  VT_ERR_OUT_OF_MEMORY = 3,    /* device or host allocation failed             */
  VT_ERR_DEVICE = 4,           /* a HIP runtime call failed                    */
  VT_ERR_NO_DEVICE = 5,        /* no HIP device, or the ordinal is invalid     */
  VT_ERR_UNSUPPORTED = 6,      /* valid request this world/mode cannot serve   */
  VT_ERR_MISALIGNED = 7,       /* a device pointer is not 16-byte aligned      */
  VT_ERR_WRONG_CONTEXT = 8     /* a world used with a context it was not made on */
};

typedef struct vt_ctx vt_ctx;
typedef struct vt_world vt_world;

/* ---- camera ------------------------------------------------------------ */
/* Byte-identical to Camera.UniformData (src/engine/graphics/camera.zig:12-16)
 * and to the std140 block u_Camera (assets/shaders/camera.glsl:2-6):
 * vec4 C_position @0, mat4 C_view @16 (column-major), float fov @80.     */
typedef struct vt_camera {
  float position[4]; /* C_position, xyz used                          */
  float view[16];    /* C_view, column j at view[4j..4j+3]            */
  float fov;         /* horizontal field of view in radians           */
  float _pad[3];
} vt_camera;

/* ---- render parameters ------------------------------------------------- */
/* flags */
#define VT_FLAG_GBUF_UNORM8 (1u << 0) /* G-buffer colour/normal/illumination  */
                                      /* stored through rgba8 clamp+round,    */
                                      /* as gbuffer.zig:10-13 declares them   */
#define VT_FLAG_LAMBERT (1u << 1)     /* shaded = albedo*(ambient+max(n.l,0)*vis) */
                                      /* instead of the blit composite; l is  */
                                      /* sun/|sun|, so a sun whose binary32   */
                                      /* squared length is 0 or overflows is  */
                                      /* VT_ERR_INVALID_ARGUMENT              */
#define VT_FLAG_ALBEDO_COMPACT (1u << 2) /* albedo / its gradient hold occupied */
                                      /* voxels only, in the order of the     */
                                      /* world's compact index                */
#define VT_FLAG_GRAD_L2 (1u << 3)     /* with LAMBERT: also accumulate the    */
                                      /* gradients of sum |shaded - target|^2 */
                                      /* (frame.target) into grad_albedo,     */
                                      /* grad_sun and loss (recompute mode)   */
#define VT_FLAG_PRESENT (1u << 4)     /* shaded = the presented colour: the   */
                                      /* composite through the crosshair and  */
                                      /* vignette of blit.fragment.glsl:32-35 */
                                      /* (not with LAMBERT); frame.display,   */
                                      /* when set, gets it as RGBA8           */
/* Free-region test (flat traversal of dense worlds; brickmap worlds with a
 * power-of-two dim).  The library keeps, per world, a bitmap of the 16^3
 * regions from which no ray of a given direction octant can meet a nonempty
 * block (every region ahead of it on all three axes is empty: traceMap's
 * blocks only move in the travel direction), and ends a ray as a miss when
 * it stands in such a region.  Every output is unchanged except the step
 * counts, so the test is never made for a frame that stores them
 * (frame.steps).  Statistics (frame.stats) count traceMap's steps; with
 * VT_FLAG_FREE_MAP_ON on a dense world they count the steps executed (rays
 * ending early).  Default: made when a step cap is at least VT_FREE_MIN_CAP
 * and some octant has at least 1/VT_FREE_MIN_SHARE of the world's regions
 * free (below either, the test costs more than it saves); these flags force
 * it either way (not both).                                                */
#define VT_FLAG_FREE_MAP_ON (1u << 5)
#define VT_FLAG_FREE_MAP_OFF (1u << 6)
#define VT_FREE_MIN_CAP 512
#define VT_FREE_MIN_SHARE 16

/* traversal storage variants for dense worlds (identical hits by contract) */
#define VT_TRAVERSAL_FLAT 0u     /* one uint8 gather per executed step       */
#define VT_TRAVERSAL_BITMASK 1u  /* 2-level uint64 occupancy: 4^3 bricks,    */
                                 /* one bit each in 16^3-region words        */
#define VT_TRAVERSAL_BITMASK8 2u /* 2-level uint64 occupancy: 8^3 bricks,    */
                                 /* one bit each in 32^3-region words        */

typedef struct vt_render_params {
  uint32_t width, height;    /* image size in pixels                        */
  int32_t max_steps_primary; /* traceMap cap, primary.comp.glsl:43 uses 192 */
  int32_t max_steps_shadow;  /* traceMap cap, secondary.comp.glsl:41 uses 48 */
  float sun[3];              /* SUN_DIR (camera.glsl:8): shadow-ray direction, */
                             /* used unnormalised; also the sky's sun         */
  uint32_t flags;            /* VT_FLAG_*                                     */
  uint32_t traversal;        /* VT_TRAVERSAL_*                                */
  float ambient;             /* VT_FLAG_LAMBERT only                          */
  const float* albedo;       /* VT_FLAG_LAMBERT only: device float RGB per    */
                             /* voxel: dense [dim^3*3], index x+dim*(y+dim*z), */
                             /* or compact [occupied*3] (VT_FLAG_ALBEDO_COMPACT) */
  float* grad_albedo;        /* VT_FLAG_GRAD_L2: += dL/dalbedo, albedo layout */
  float* grad_sun;           /* VT_FLAG_GRAD_L2: += dL/dsun, device float[3]  */
  float* loss;               /* VT_FLAG_GRAD_L2: += L, device float[1]        */
} vt_render_params;

/* ---- outputs ------------------------------------------------------------ */
/* Per-pixel hit record (12 bytes).  voxel = linear block index
 * x + dim*(y + dim*z) of the hit block (0xFFFFFFFF on a miss); t = distance
 * from C_position to hit_pos/8; info packs:
 *   bits 0-2  face id 1..6 (map.glsl:119-125; 0 = miss)
 *   bit  3    shadow ray occluded (terrain or entity)
 *   bits 4-12 sub-voxel index x+8y+64z inside the block (brickmap worlds)
 *   bits 16-23 material id (dense worlds)                                  */
typedef struct vt_hit_record {
  uint32_t voxel;
  float t;
  uint32_t info;
} vt_hit_record;

#define VT_INFO_FACE(info) ((info)&7u)
#define VT_INFO_SHADOWED(info) (((info) >> 3) & 1u)
#define VT_INFO_SUBVOXEL(info) (((info) >> 4) & 511u)
#define VT_INFO_MATERIAL(info) (((info) >> 16) & 255u)

/* Statistics accumulated when vt_frame.stats != NULL (device uint64[8]). */
enum {
  VT_STAT_PRIMARY_STEPS = 0, /* executed traceMap iterations, primary rays  */
  VT_STAT_SHADOW_STEPS = 1,  /* executed traceMap iterations, shadow rays   */
  VT_STAT_PRIMARY_HITS = 2,  /* = number of shadow rays traced              */
  VT_STAT_OCCLUDED = 3,      /* shadow rays that found an occluder          */
  VT_STAT_L1_FETCHES = 4,    /* bitmask traversal: 16^3 region words loaded; */
                             /* brickmap worlds: chunk-table reads; flat    */
                             /* traversal with VT_FLAG_FREE_MAP_ON: steps   */
                             /* taken without a lookup (a primary ray's     */
                             /* steps before its line can reach the box of  */
                             /* the nonempty voxels)                        */
  VT_STAT_L0_FETCHES = 5,    /* bitmask traversal: 4^3 brick words loaded;  */
                             /* brickmap worlds: block-pool + atlas reads   */
  VT_STAT_LEAF_READS = 6,    /* bitmask traversal: material bytes read at   */
                             /* primary hits (flat: 0, its gathers are one  */
                             /* byte per executed step with a lookup)       */
  VT_STAT_ENTITY_HITS = 7,   /* shadow rays occluded by traceEntities       */
  VT_STAT_COUNT = 8
};

/* Caller-owned device buffers; any may be NULL.  Images are row-major,
 * row y = pixel y (the reference's gl_GlobalInvocationID.y), 16-byte
 * aligned. */
typedef struct vt_frame {
  float* shaded;         /* W*H*4: RGBA32F composite (or Lambert) result      */
  vt_hit_record* hits;   /* W*H records                                       */
  float* color;          /* W*H*4: G-buffer image 0 (frameColor)              */
  float* normal;         /* W*H*4: G-buffer image 1 (frameNormal)             */
  float* position;       /* W*H*4: G-buffer image 2 (framePosition)           */
  float* illumination;   /* W*H*4: G-buffer image 3 (frameIllumination)       */
  uint32_t* steps;       /* W*H: primary steps | shadow steps << 16           */
  unsigned long long* stats; /* VT_STAT_COUNT counters, accumulated (+=)      */
  const float* target;   /* W*H*4 input: VT_FLAG_GRAD_L2 target image (RGB)   */
  uint32_t* display;     /* W*H RGBA8 (R in the low byte): VT_FLAG_PRESENT    */
                         /* colour stored as an rgba8 framebuffer does        */
                         /* (clamp to [0,1], round to nearest)                */
  uint32_t* schedule;    /* VT_SCHEDULE_WORDS(W, H) words, zero-filled before */
                         /* its first render: the tile schedule (below)       */
} vt_frame;

/* Tile schedule (vt_frame.schedule; vt_render only, the batch ignores it).
 * No counterpart in the reference, which dispatches its compute shaders over
 * the image in a fixed order (game.zig:241-245).  vt_render launches a
 * frame's 256-pixel blocks, more than the GPU holds at once; in image order
 * the costly blocks (rays that skim the terrain toward the horizon: several
 * times the steps of the others) start mid-launch and run on alone at its
 * end.  With a schedule buffer each render measures the time of every block
 * and the next render into the same frame launches the blocks most costly
 * first, from that record.  The images are the same in either
 * order: only which launch slot renders which block changes.  The buffer is
 * private to one frame (W x H); zero-fill it before its first render (and
 * again before using it for another image size: a buffer last used at another
 * size is detected and restarted, but only a zero-filled one is guaranteed
 * to be).  Renders into one frame must not overlap, as for its images.  The
 * first render into a frame, and renders into frames without a schedule, use
 * the fixed order. */
#define VT_SCHEDULE_BLOCKS(W, H) ((((W) + 15u) / 16u) * (((H) + 7u) / 8u))
#define VT_SCHEDULE_WORDS(W, H) \
  (512u + 3u * (4096u + 128u * ((VT_SCHEDULE_BLOCKS(W, H) + 7u) / 8u)))

/* ---- context ------------------------------------------------------------ */
/* Replaces the GL context made current in gfx.init (graphics.zig:60-75). */
int vt_create(vt_ctx** out, int device);
int vt_destroy(vt_ctx* ctx);
/* Blocks until all work the library enqueued on `stream` finished. */
int vt_sync(vt_ctx* ctx, void* stream);
const char* vt_status_string(int status);
/* Detail of the last failure on the calling thread (validation message or
 * HIP error string); "" when none. */
const char* vt_last_error(void);
/* Number of compute units of the context's device (for launch sizing). */
int vt_device_info(const vt_ctx* ctx, int* compute_units, char* name, size_t name_len);

/* ---- worlds -------------------------------------------------------------- */
/* Dense uint8 world: grid[x + dim*(y + dim*z)], 0 = empty, k>0 = material k.
 * One grid voxel plays one reference block whose atlas slot is a solid 8^3
 * cube of colour palette[k] (RGBA8, R in the low byte), so traceMap never
 * descends to the fine level.  dim must be a power of two, 16..1024
 * (MAP_DIMENSION is 512 in map.glsl:2), and palette[1..255] nonzero (a zero
 * texel reads as empty).  On the device the voxels are stored in Morton
 * order with 4^3-brick / 16^3-region occupancy bitmasks.
 * Replaces VoxelBrickmap(512,8) + VoxelModelAtlas (src/engine/voxel.zig:25-131)
 * for the config workloads; also builds the 2-level occupancy bitmasks. */
int vt_world_create_dense_u8(vt_ctx* ctx, const uint8_t* grid, uint32_t dim,
                             const uint32_t palette_rgba8[256], vt_world** out);

/* Reference layout (drop-in): chunk table chunks[(dim/8)^3] (0 = empty chunk,
 * else pool block index + 1), block pool pool[pool_words] (512 u32 per block,
 * in-block index x + 8y + 64z), atlas_rgba8[256^3] RGBA8 texels, index
 * x + 256*(y + 256*z).  Exactly the buffers VoxelBrickmap.bind and
 * VoxelModelAtlas.bind expose (voxel.zig:77-80,129-131); dim = MAP_DIMENSION
 * (map.glsl:2), a multiple of 8.                                            */
int vt_world_create_brickmap(vt_ctx* ctx, uint32_t dim, const uint32_t* chunks,
                             const uint32_t* pool, size_t pool_words,
                             const uint32_t* atlas_rgba8, vt_world** out);
/* MagicaVoxel model files -> the model atlas (host side, no device work).
 * vt_atlas_load_vox replaces VoxelModelAtlas.load_block_model
 * (voxel.zig:94-127): every model of the file (SIZE/XYZI pairs in file
 * order) takes the next slot k = *next_slot, its 8^3 box at
 * ((k%32)*8, (k/32%32)*8, (k/1024%1024)*8) of atlas_rgba8 (256^3 texels,
 * index x + 256*(y + 256*z)) is cleared and voxel (x, y, z, colour c) is
 * written at box offset (x, z, y) with the file's RGBA palette entry c-1;
 * *next_slot advances by the model count.  Files without an RGBA chunk are
 * rejected (VT_ERR_UNSUPPORTED).  vt_vox_info counts the models.          */
int vt_vox_info(const uint8_t* bytes, size_t len, uint32_t* n_models, uint32_t* version);
int vt_atlas_load_vox(const uint8_t* bytes, size_t len, uint32_t* atlas_rgba8,
                      uint32_t* next_slot);

/* ---- world edits, reads and the centre-ray pick ---------------------------- */
/* Edits: n voxels at block coordinates xyz[3i..3i+2] (each < dim) get
 * values[i], with the meaning of calling VoxelBrickmap.set (voxel.zig:58-61)
 * once per edit in order: brickmap worlds take reference block words
 * (Voxel: ty | is_solid << 28, voxel.zig:7-19) and allocate a chunk's pool
 * block on its first edit (voxel.zig:47-56; the pool grows by doubling,
 * gpu_block_allocator.zig:20-30); dense worlds take materials 0..255 and
 * keep their occupancy masks (and the compact index, when built) current.
 * Of several edits of one voxel the last wins.  The call returns after the
 * edits are applied on `stream`; renders on other streams must be ordered
 * after it by the caller.                                                   */
int vt_world_set_voxels(vt_ctx* ctx, vt_world* world, const uint32_t* xyz,
                        const uint32_t* values, uint32_t n, void* stream);
/* Reads: values[i] = map_getVoxel of block xyz[3i..] (map.glsl:38-47), the
 * block word for brickmap worlds (0 where the chunk is unallocated), the
 * material for dense worlds.  Host arrays; returns when values are written. */
int vt_world_get_voxels(vt_ctx* ctx, const vt_world* world, const uint32_t* xyz, uint32_t n,
                        uint32_t* values, void* stream);
/* The chunk table and block pool of a brickmap world as the CPU reads them
 * through VoxelBrickmap's persistent mappings (voxel.zig:32-39,63-70):
 * chunks[(dim/8)^3] and pool[0 .. *pool_words), 512 words per allocated
 * block.  chunks or pool may be NULL; *pool_words is always set; a
 * pool_capacity_words below the allocated words is an error
 * (VT_ERR_INVALID_ARGUMENT).  Returns when the outputs are written.       */
/* The free-region map of a world (VT_FLAG_FREE_MAP_ON above): 8 octant
 * bitmaps of *words_per_octant uint32 words each, octant o = (x negative) |
 * (y negative) << 1 | (z negative) << 2, bit r = region with Morton code r
 * (regions of 16^3 blocks, max(dim/16, 1) per axis) free for that octant.
 * words may be NULL (only *words_per_octant is set); otherwise it holds
 * 8 * *words_per_octant words.  *valid = 1 when the map is current (0 for
 * worlds without one, or after a failed rebuild).  No counterpart in the
 * reference; for checking the map against the world.  Returns when the
 * words are written.                                                       */
int vt_world_read_free_map(vt_ctx* ctx, const vt_world* world, uint32_t* words,
                           size_t capacity_words, uint32_t* words_per_octant, int* valid,
                           void* stream);
int vt_world_read_brickmap(vt_ctx* ctx, const vt_world* world, uint32_t* chunks,
                           uint32_t* pool, size_t pool_capacity_words, size_t* pool_words,
                           void* stream);

/* The reference world generator, procgen(dim, &voxels, offset_x, offset_y)
 * (procgen.zig:6-70, called by Game.init as procgen(512, &voxels, 0, 0),
 * game.zig:96-97), on an empty brickmap world (no pool blocks yet; dim a
 * multiple of 8, at least 16): a stone base layer, FastNoiseLite fBm
 * terrain, decorations and trees, with the chunk table and pool equal word
 * for word to what the generator's VoxelBrickmap.set calls leave.  A tree or
 * terrain column that would write outside the map (small dims; the
 * generator would index past its chunk table) is VT_ERR_INVALID_ARGUMENT and
 * leaves the world unchanged.  Returns when the world is written.          */
int vt_world_procgen(vt_ctx* ctx, vt_world* world, float offset_x, float offset_y,
                     void* stream);

/* vt_world_get_voxels, vt_pick and vt_world_read_brickmap allocate no device
 * memory: they stage through a buffer each vt_ctx allocates once.  A failed
 * vt_world_set_voxels or vt_world_procgen leaves the world as it was.      */
/* terrain_edit.comp.glsl:11-24: the camera's centre ray (rayUV = 0) through
 * the AABB entry and traceMap with max_steps (64 in the shader).  place /
 * remove are the block cells the shader's (disabled) edit body would write:
 * ivec3(hit_pos + normal) >> 3 and ivec3(hit_pos) >> 3.                     */
typedef struct vt_pick_result {
  uint32_t data;      /* HitInfo.data: RGBA8 of the hit sub-voxel, 0 = miss */
  uint32_t face;      /* 1..6 (normals[face-1], map.glsl:72-79), 0 = miss   */
  uint32_t steps;     /* traceMap iterations executed                       */
  uint32_t _pad;
  float hit_pos[3];   /* HitInfo.hit_pos in sub-voxel units; (-1,-1,-1) miss */
  float normal[3];    /* HitInfo.normal; (0,0,0) on a miss                   */
  int32_t place[3];
  int32_t remove[3];
} vt_pick_result;
int vt_pick(vt_ctx* ctx, const vt_world* world, const vt_camera* camera, int32_t max_steps,
            vt_pick_result* result, void* stream);

int vt_world_destroy(vt_world* world);
/* dim and device bytes held by the world. */
int vt_world_info(const vt_world* world, uint32_t* dim, size_t* device_bytes);

/* ---- render ------------------------------------------------------------ */
/* Primary pass (primary.comp.glsl:23-69) + shadow pass
 * (secondary.comp.glsl:18-50) + blit composite formula
 * (blit.fragment.glsl:30-31), fused into one launch; replaces the
 * dispatches of Game.render (game.zig:241-248).                           */
int vt_render(vt_ctx* ctx, const vt_world* world, const vt_camera* camera,
              const vt_render_params* params, const vt_frame* frame, void* stream);

/* Whether vt_render with these arguments makes the free-region test
 * (VT_FLAG_FREE_MAP_ON above): *used = 1 when its rays end early in free
 * regions, 0 when every ray runs to its end.  No counterpart in the
 * reference, whose traceMap runs every ray to its end; for callers that
 * need to know whether a frame's statistics are traceMap's step counts.   */
int vt_free_map_used(const vt_world* world, const vt_render_params* params,
                     const vt_frame* frame, int* used);

/* ---- present pass ---------------------------------------------------------- */
/* The blit pass after the composite (blit.fragment.glsl:32-35), as a
 * stand-alone streaming kernel over a composite image: for pixel (x, y) at
 * texPos = ((x+0.5)/W, (y+0.5)/H) (pixel centres of the full-screen quad,
 * blit.vertex.glsl:5-15), colour = mix(c, (1,1,1,0.4), 0.5) within 0.002 of
 * the centre (crosshair), then * pow(u(1-u) * v(1-v) * 15, 0.6 * 0.3)
 * (vignette), with (u, v) = texPos.  out (RGBA32F) and/or display (RGBA8) may be NULL, not both.
 * vt_render with VT_FLAG_PRESENT does the same inside the render kernel. */
int vt_present(vt_ctx* ctx, const float* composite, uint32_t width, uint32_t height, float* out,
               uint32_t* display, void* stream);

/* ---- differentiable Lambert shading ---------------------------------------- */
/* Project-defined (no reference counterpart: the reference shades with the
 * blit composite).  shaded = albedo[voxel] * (ambient + max(n . sun/|sun|, 0)
 * * vis) on hit pixels, 0 on sky pixels; n = the hit face's normal, vis = 0
 * when the shadow ray was occluded.  The discrete hit structure (voxel, face,
 * shadow bit) is held fixed: gradients are those of the shading alone.      */

/* Builds (once) the compact albedo index of a dense world: occupied voxels
 * numbered in Morton order, through an exclusive prefix sum over the
 * popcounts of the 4^3-brick occupancy words.  Returns the occupied count. */
int vt_world_build_compact_index(vt_ctx* ctx, vt_world* world, uint64_t* n_occupied);

/* upstream kinds for vt_lambert_backward */
#define VT_UPSTREAM_GRADIENT 0u  /* upstream = dL/dshaded (RGBA32F, rgb used)   */
#define VT_UPSTREAM_L2_TARGET 1u /* upstream = target T: L = sum |shaded-T|^2   */

/* Shading from stored hit records (the forward of the backward below):
 * out[i] = shaded of hits[i].  Uses params->albedo, sun, ambient, flags
 * (VT_FLAG_ALBEDO_COMPACT).  npix records, out RGBA32F. */
int vt_lambert_shade(vt_ctx* ctx, const vt_world* world, const vt_hit_record* hits,
                     uint32_t npix, const vt_render_params* params, float* out, void* stream);
/* Gradients from stored hit records: grad_albedo += dL/dalbedo (albedo
 * layout), grad_sun[3] += dL/dsun, loss[0] += L (VT_UPSTREAM_L2_TARGET only;
 * may be NULL).  Float atomics: sums are order-dependent to rounding. */
int vt_lambert_backward(vt_ctx* ctx, const vt_world* world, const vt_hit_record* hits,
                        uint32_t npix, const vt_render_params* params, const float* upstream,
                        uint32_t upstream_kind, float* grad_albedo, float* grad_sun,
                        float* loss, void* stream);

/* ---- batched rendering --------------------------------------------------- */
/* Many views of uneven resolution in ONE persistent launch: waves pull
 * 64-pixel tiles of all views from per-XCD work queues.  No reference
 * counterpart (the engine renders one view per frame, game.zig:232-256); it
 * runs the same per-pixel path as vt_render, so every view's outputs are
 * bit-identical to a vt_render of that view.  Dense worlds only.           */
typedef struct vt_view {
  vt_camera camera;
  uint32_t width, height;
  uint32_t _pad[2];
  vt_frame frame; /* this view's outputs; frame.stats and frame.schedule */
                  /* are ignored                                        */
} vt_view;

/* wave tile shapes (log2 of the tile width; 64 pixels per tile) */
#define VT_TILE_8X8 3u
#define VT_TILE_16X4 4u
#define VT_TILE_32X2 5u

typedef struct vt_tuning {
  uint32_t tile;          /* VT_TILE_*                                      */
  uint32_t chunk;         /* tiles a wave takes per queue grab, 1..64       */
  uint32_t traversal;     /* VT_TRAVERSAL_* (overrides params->traversal):  */
                          /* the storage walked and its brick size         */
  uint32_t blocks_per_cu; /* persistent 256-thread blocks per CU, 1..16     */
} vt_tuning;

typedef struct vt_batch vt_batch;
/* Copies the view table to the device (a fixed argument buffer, so the
 * launch can be captured in a graph and replayed).  params->width/height are
 * ignored (each view has its own); stats may be NULL (device uint64[8]).    */
int vt_batch_create(vt_ctx* ctx, const vt_world* world, const vt_view* views, uint32_t n_views,
                    const vt_render_params* params, const vt_tuning* tuning,
                    unsigned long long* stats, vt_batch** out);
/* Enqueues the queue reset and the persistent launch on `stream`. */
int vt_batch_launch(vt_batch* batch, void* stream);
/* Captures reset + launch into a hipGraph (on a stream the library owns);
 * vt_batch_graph_launch then replays it on any stream.  Calls that overlap
 * on one batch race on its queue counters. */
int vt_batch_capture_graph(vt_batch* batch);
int vt_batch_graph_launch(vt_batch* batch, void* stream);
int vt_batch_destroy(vt_batch* batch);
/* The autotuner: times every tile shape x chunk {1,2,4,8} x traversal
 * (flat; bitmask with 4^3 or with 8^3 bricks) on these views (`reps` timed launches each after one
 * warm-up) and returns the fastest tuning and its mean launch time.  Every
 * timed launch writes the views' frames (their contents are those of a
 * render when it returns); VT_FLAG_GRAD_L2 is rejected
 * (VT_ERR_INVALID_ARGUMENT), since each launch would add into the caller's
 * gradient accumulators -- tune with VT_FLAG_LAMBERT, the same kernel
 * shape without the accumulation. */
int vt_batch_autotune(vt_ctx* ctx, const vt_world* world, const vt_view* views,
                      uint32_t n_views, const vt_render_params* params, int reps,
                      vt_tuning* best, float* best_ms);

/* ---- batched rendering without allocation ------------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Renders the views like vt_batch_create + vt_batch_launch (same
 * kernels, same outputs bit for bit) but allocates nothing: the view table
 * and the work-queue counters live in a fixed device array of the library
 * (about 45 KiB per GPU, held until the process ends), whose address each
 * vt_create takes once per device.  The view records are handed over as
 * kernel arguments, so `views` may be freed when the call returns.  At most
 * VT_NOALLOC_MAX_VIEWS views (else VT_ERR_INVALID_ARGUMENT).  All calls on
 * one GPU share that array: calls that overlap on one GPU (other streams,
 * other threads or contexts) race; order them on one stream.             */
#define VT_NOALLOC_MAX_VIEWS 256u
int vt_batch_render_noalloc(vt_ctx* ctx, const vt_world* world, const vt_view* views,
                            uint32_t n_views, const vt_render_params* params,
                            const vt_tuning* tuning, unsigned long long* stats, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* VT_H */
