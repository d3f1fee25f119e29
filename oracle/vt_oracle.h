/*
 * vt_oracle.h -- CPU restatement of the reference voxel ray-traversal path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the test
 * oracle (or the timed CPU baseline); the product library never links it.
 *
 * It restates, in plain C99 on IEEE-754 binary32 (compiled with
 * -ffp-contract=off, so every multiply and add rounds separately):
 *   primary raygen            assets/shaders/primary.comp.glsl:31-39
 *   intersectAABB + entry     assets/shaders/map.glsl:21-29, primary.comp.glsl:42-43
 *   traceMap (REF DDA)        assets/shaders/map.glsl:83-168
 *   map_getChunkFlags/Voxel   assets/shaders/map.glsl:31-47
 *   map_getSubVoxel           assets/shaders/map.glsl:57-60
 *   traceEntities             assets/shaders/map.glsl:172-201
 *   primary stores + sky      assets/shaders/primary.comp.glsl:58-68, camera.glsl:11-19
 *   shadow pass               assets/shaders/secondary.comp.glsl:18-50
 *   blit composite formula    assets/shaders/blit.fragment.glsl:30-31
 * plus the project-defined Lambert x shadow shading and its FP64 gradient.
 *
 * Operations GLSL leaves to the implementation are fixed as follows (the
 * HIP kernels follow the same choices): a/b and sqrt correctly rounded,
 * normalize(v) = v / sqrt(dot(v,v)) with the dot summed left to right,
 * mat4*vec4 summed column 0..3 left to right, min/max return the non-NaN
 * operand, uvec3(float) of a value in (-1,0) is 0, fract(x) = x - floor(x),
 * tan(fov/2) = (float)tan((double)(fov/2)) evaluated once per view.
 */
#ifndef VT_ORACLE_H
#define VT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_camera {
  float position[4];
  float view[16];
  float fov;
  float _pad[3];
} or_camera;

enum { OR_WORLD_DENSE = 0, OR_WORLD_BRICKMAP = 1 };

typedef struct or_world {
  int kind;
  uint32_t dim; /* MAP_DIMENSION: blocks per axis */
  /* dense: grid[x + dim*(y + dim*z)], palette[256] RGBA8 */
  const uint8_t* grid;
  const uint32_t* palette;
  /* brickmap: chunks[(dim/8)^3], pool[], atlas[256^3] */
  const uint32_t* chunks;
  const uint32_t* pool;
  uint64_t pool_words;
  const uint32_t* atlas;
} or_world;

/* HitInfo (map.glsl:62-70) plus bookkeeping. */
typedef struct or_hit {
  uint32_t data;     /* packUnorm4x8 texel; 0 = miss                  */
  float hit_pos[3];  /* gridsCoords + withinGridCoords (x8 units)     */
  float normal[3];
  int face;          /* 1..6, 0 on miss                               */
  int32_t block[3];  /* pos >> 3 at the hit                           */
  uint32_t sub;      /* pos % 8 packed x + 8y + 64z                   */
  int steps;         /* executed loop iterations                      */
  float tie_margin;  /* min over DDA steps of (2nd smallest - smallest t) */
  uint32_t lookups;  /* in-bounds block lookups (dense worlds)        */
} or_hit;

/* tan(fov/2) as evaluated once per view. */
float or_tan_half_fov(float fov);
/* Primary ray of pixel (px,py) in a W x H image (primary.comp.glsl:31-39). */
void or_raygen(const or_camera* cam, float tan_half, uint32_t W, uint32_t H,
               uint32_t px, uint32_t py, float o[3], float d[3]);
/* intersectAABB (map.glsl:21-29). */
void or_intersect_aabb(const float o[3], const float d[3], const float bmin[3],
                       const float bmax[3], float* t_near, float* t_far);
/* Primary entry origin o + d*max(tNear,0) - EPSILON (primary.comp.glsl:42-43). */
void or_entry_origin(const float o[3], const float d[3], uint32_t dim, float out[3]);
/* traceMap (map.glsl:83-168).  path (optional, cap entries of 3 ints)
 * receives `pos` of every executed lookup. Returns hit->data != 0. */
int or_trace_map(const or_world* w, const float o[3], const float d[3], int max_steps,
                 or_hit* hit, int32_t* path, int path_cap);
/* traceEntities occlusion test (map.glsl:172-201): 1 if any entity box hits. */
int or_trace_entities(const float o[3], const float d[3], float max_distance);
/* SkyDome2 (camera.glsl:11-19): rgb. */
void or_sky(const float sun[3], const float rd[3], float rgb[3]);

/* Per-pixel outputs of a render, any pointer may be NULL (same layouts as
 * vt_frame in include/vt.h). */
typedef struct or_frame {
  float* shaded;     /* W*H*4 */
  uint32_t* hits;    /* W*H*3 words: voxel, t bits, info */
  float* color;      /* W*H*4 */
  float* normal;     /* W*H*4 */
  float* position;   /* W*H*4 */
  float* illumination; /* W*H*4 */
  uint32_t* steps;   /* W*H */
  float* tie_margin; /* W*H: min of primary and shadow margins */
  uint64_t* stats;   /* 8 counters, accumulated */
  uint32_t* display; /* W*H RGBA8 of the presented colour (flags bit4) */
} or_frame;

typedef struct or_params {
  uint32_t width, height;
  int32_t max_steps_primary, max_steps_shadow;
  float sun[3];
  uint32_t flags;     /* bit0 GBUF_UNORM8, bit1 LAMBERT, bit4 PRESENT */
  uint32_t traversal; /* 0 flat, 1 / 2 bitmask with 4^3 / 8^3 bricks (dense worlds:
                         memory counters only) */
  float ambient;
  const float* albedo; /* LAMBERT: dim^3*3 floats */
} or_params;

/* Render rows [y0, y1) of the frame with `threads` POSIX threads. */
void or_render(const or_world* w, const or_camera* cam, const or_params* p,
               const or_frame* f, uint32_t y0, uint32_t y1, int threads);

/* terrain_edit.comp.glsl:11-24: the centre ray (rayUV = 0) through the AABB
 * entry, traceMap(.., max_steps); place / remove = the edit cells
 * ivec3(hit_pos + normal) >> 3 and ivec3(hit_pos) >> 3. */
typedef struct or_pick_result {
  uint32_t data, face, steps, _pad;
  float hit_pos[3], normal[3];
  int32_t place[3], remove[3];
} or_pick_result;
void or_pick(const or_world* w, const or_camera* cam, int max_steps, or_pick_result* out);

/* Present pass of one composite image (blit.fragment.glsl:32-35): crosshair
 * and vignette at texPos = ((x+0.5)/W, (y+0.5)/H); out RGBA32F and/or
 * display RGBA8 (clamp, round to nearest, R in the low byte), either NULL. */
void or_present(const float* composite, uint32_t W, uint32_t H, float* out, uint32_t* display);

/* Lambert x visibility shading from hit records (project-defined):
 * rgb = albedo[voxel] * (ambient + max(dot(n, sun/|sun|), 0) * vis),
 * sky pixels 0. */
void or_shade_lambert(const uint32_t* hits, uint32_t npix, const float* albedo,
                      const float sun[3], float ambient, float* rgba_out);
/* FP64 gradient of L = sum (I - target)^2 over rgb of all pixels with the
 * hit structure fixed: dL/dalbedo (dense dim^3*3 doubles, accumulated) and
 * dL/dsun (3 doubles, accumulated).  Returns L. */
double or_lambert_grad64(const uint32_t* hits, uint32_t npix, const float* albedo,
                         const float sun[3], float ambient, const float* target_rgba,
                         double* dalbedo, double* dsun);

/* ---- reference world generator (vt_procgen_oracle.c) ---- */
/* FastNoiseLite gradient table (256 floats, pairs gx, gy). */
void or_fnl_gradients(float out[256]);
/* _fnlSingleSimplex2D on skewed coordinates. */
float or_fnl_simplex2(int seed, float x, float y);
/* FnlGenerator{.fractal_type = .fbm}.noise2 (procgen.zig:7,22). */
float or_fnl_fbm2(float x, float y);
/* procgen.zig:22-23: terrain height of column (x, z). */
uint32_t or_procgen_height(uint32_t dim, float offset_x, float offset_y, uint32_t x, uint32_t z);
/* procgen.zig:6-70 into an empty VoxelBrickmap(dim, 8); 0 = ok, 1 = pool
 * too small, 2 = a write outside the map, 3 = bad dim. */
int or_procgen(uint32_t dim, float offset_x, float offset_y, uint32_t* chunks, uint32_t* pool,
               uint64_t pool_capacity_words, uint64_t* pool_words);

#ifdef __cplusplus
}
#endif
#endif
