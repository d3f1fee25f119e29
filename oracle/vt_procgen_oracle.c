/*
 * vt_procgen_oracle.c -- CPU restatement of the reference world generator.
 *
 * TEST INFRASTRUCTURE ONLY (the test oracle and the timed CPU baseline); the
 * product library never links or calls it.  Built into liboracle.so by
 * oracle/Makefile (gcc -O2 -ffp-contract=off: every binary32 operation
 * rounds to nearest-even once).
 *
 * or_procgen follows src/procgen.zig:6-70 line by line, writing through a
 * literal VoxelBrickmap(dim, 8) (src/engine/voxel.zig:45-70: a chunk's pool
 * block is allocated on its first set(), index = allocation count,
 * gpu_block_allocator.zig:19-30), in the generator's own order.
 *
 * The heights come from znoise's FnlGenerator{ .fractal_type = .fbm }.noise2
 * (procgen.zig:7,22).  znoise (zig-gamedev/znoise @ 96f9458c, 0.3.0-dev, a
 * dependency in build.zig.zon that is not vendored in the reference) wraps
 * the C edition of FastNoiseLite (Jordan Peck, MIT); its generator defaults
 * are seed 1337, frequency 0.01, OpenSimplex2, 3 octaves, lacunarity 2,
 * gain 0.5, weighted strength 0.  The noise below restates FastNoiseLite's
 * published algorithm: fnlGetNoise2D -> coordinate transform (frequency,
 * then the simplex skew) -> fBm over 2D simplex noise with the hashed
 * 24+8-direction gradient table.  How it is pinned (lattice zeros, an
 * independent vertex-sum evaluation in float64, the gradient directions) is
 * in tests/test_procgen_oracle.py.
 */
#include "vt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* FastNoiseLite, 2D OpenSimplex2 + fBm                                */
/* ------------------------------------------------------------------ */

#define FNL_PRIME_X 501125321u
#define FNL_PRIME_Y 1136930381u

/* Gradient pairs (gx, gy) = (sin a, cos a): pairs 0..119 run through the
 * 24 directions a = 7.5 + 15 m degrees (m = pair % 24) five times, pairs
 * 120..127 are the 8 directions a = 22.5 + 45 j degrees.  Computed here in
 * double from the angles and rounded once to binary32. */
static float g_grad2[256];
static int g_grad2_ready;

static void grad_init(void) {
  if (g_grad2_ready) return;
  const double deg = 3.14159265358979323846 / 180.0;
  for (int p = 0; p < 128; p++) {
    double a = p < 120 ? 7.5 + 15.0 * (p % 24) : 22.5 + 45.0 * (p - 120);
    g_grad2[2 * p] = (float)sin(a * deg);
    g_grad2[2 * p + 1] = (float)cos(a * deg);
  }
  g_grad2_ready = 1;
}

void or_fnl_gradients(float out[256]) {
  grad_init();
  memcpy(out, g_grad2, sizeof g_grad2);
}

/* _fnlFastFloor: (int)f for f >= 0, else (int)f - 1 (so an exact negative
 * integer maps one below its floor, as in the library). */
static inline int fnl_floor(float f) { return f >= 0 ? (int)f : (int)f - 1; }

/* _fnlGradCoord2D: hash = (seed ^ xPrimed ^ yPrimed) * 0x27d4eb2d (int,
 * wrapping), hash ^= hash >> 15, pair (hash & 254) / 2. */
static inline float fnl_grad(int seed, uint32_t xp, uint32_t yp, float xd, float yd) {
  uint32_t h = ((uint32_t)seed ^ xp ^ yp) * 0x27d4eb2du;
  h ^= (uint32_t)((int32_t)h >> 15);
  h &= 254u;
  return xd * g_grad2[h] + yd * g_grad2[h | 1u];
}

/* _fnlSingleSimplex2D on already skewed coordinates. */
float or_fnl_simplex2(int seed, float x, float y) {
  grad_init();
  const float SQRT3 = 1.7320508075688772935274463415059f;
  const float G2 = (3 - SQRT3) / 6;
  int i = fnl_floor(x), j = fnl_floor(y);
  float xi = x - (float)i, yi = y - (float)j;
  float t = (xi + yi) * G2;
  float x0 = xi - t, y0 = yi - t;
  uint32_t ip = (uint32_t)i * FNL_PRIME_X, jp = (uint32_t)j * FNL_PRIME_Y;
  float n0, n1, n2;
  float a = 0.5f - x0 * x0 - y0 * y0;
  if (a <= 0)
    n0 = 0;
  else
    n0 = (a * a) * (a * a) * fnl_grad(seed, ip, jp, x0, y0);
  const float K1 = (float)(2 * (1 - 2 * G2) * (1 / G2 - 2));
  const float K2 = (float)(-2 * (1 - 2 * G2) * (1 - 2 * G2));
  float c = K1 * t + (K2 + a);
  if (c <= 0) {
    n2 = 0;
  } else {
    float x2 = x0 + (2 * G2 - 1), y2 = y0 + (2 * G2 - 1);
    n2 = (c * c) * (c * c) * fnl_grad(seed, ip + FNL_PRIME_X, jp + FNL_PRIME_Y, x2, y2);
  }
  if (y0 > x0) {
    float x1 = x0 + G2, y1 = y0 + (G2 - 1);
    float b = 0.5f - x1 * x1 - y1 * y1;
    if (b <= 0)
      n1 = 0;
    else
      n1 = (b * b) * (b * b) * fnl_grad(seed, ip, jp + FNL_PRIME_Y, x1, y1);
  } else {
    float x1 = x0 + (G2 - 1), y1 = y0 + G2;
    float b = 0.5f - x1 * x1 - y1 * y1;
    if (b <= 0)
      n1 = 0;
    else
      n1 = (b * b) * (b * b) * fnl_grad(seed, ip + FNL_PRIME_X, jp, x1, y1);
  }
  return (n0 + n1 + n2) * 99.83685446303647f;
}

/* fnlGetNoise2D with the generator of procgen.zig:7. */
float or_fnl_fbm2(float x, float y) {
  /* _fnlTransformNoiseCoordinate2D: frequency, then the OpenSimplex2 skew */
  const float frequency = 0.01f;
  x *= frequency;
  y *= frequency;
  const float SQRT3 = (float)1.7320508075688772935274463415059;
  const float F2 = 0.5f * (SQRT3 - 1);
  float t = (x + y) * F2;
  x += t;
  y += t;
  /* _fnlGenFractalFBM2D */
  const int octaves = 3;
  const float lacunarity = 2.0f, gain = 0.5f, weighted_strength = 0.0f;
  float g = fabsf(gain), amp = g, amp_fractal = 1.0f; /* _fnlCalculateFractalBounding */
  for (int i = 1; i < octaves; i++) {
    amp_fractal += amp;
    amp *= g;
  }
  int seed = 1337;
  float sum = 0;
  amp = 1.0f / amp_fractal;
  for (int i = 0; i < octaves; i++) {
    float noise = or_fnl_simplex2(seed++, x, y);
    sum += noise * amp;
    float m = noise + 1 < 2 ? noise + 1 : 2;
    amp *= 1.0f + weighted_strength * (m * 0.5f - 1.0f);
    x *= lacunarity;
    y *= lacunarity;
    amp *= gain;
  }
  return sum;
}

/* procgen.zig:22-23: the terrain height of column (x, z). */
uint32_t or_procgen_height(uint32_t dim, float offset_x, float offset_y, uint32_t x,
                           uint32_t z) {
  float val = or_fnl_fbm2((offset_x + (float)x) / 10.0f, (offset_y + (float)z) / 10.0f);
  float h = val * (float)dim * 0.1f;
  return (uint32_t)(h > 0.0f ? h : 0.0f);
}

/* ------------------------------------------------------------------ */
/* VoxelBrickmap(dim, 8) and the generator                              */
/* ------------------------------------------------------------------ */

typedef struct bm {
  uint32_t dim, M;
  uint32_t* chunks;
  uint32_t* pool;
  uint64_t cap_blocks, used;
  int err;
} bm;

#define SOLID (1u << 28) /* Voxel.is_solid (voxel.zig:7-11) */

static uint32_t* bm_word(bm* w, uint64_t x, uint64_t y, uint64_t z, int alloc) {
  if (x >= w->dim || y >= w->dim || z >= w->dim) {
    if (!w->err) w->err = 2;
    return NULL;
  }
  uint32_t* e = &w->chunks[x / 8 + w->M * (y / 8 + w->M * (z / 8))];
  if (*e == 0) {
    if (!alloc) return NULL;
    if (w->used >= w->cap_blocks) { /* get_block_for_chunk -> alloc */
      if (!w->err) w->err = 1;
      return NULL;
    }
    *e = (uint32_t)(++w->used);
  }
  return &w->pool[(uint64_t)(*e - 1) * 512u + x % 8 + 8 * ((y % 8) + 8 * (z % 8))];
}

static void bm_set(bm* w, uint64_t x, uint64_t y, uint64_t z, uint32_t v) {
  uint32_t* p = bm_word(w, x, y, z, 1);
  if (p) *p = v;
}

static uint32_t bm_get(bm* w, uint64_t x, uint64_t y, uint64_t z) {
  uint32_t* p = bm_word(w, x, y, z, 0);
  return p ? *p : 0u;
}

/* util.LCG.rand (src/engine/util.zig:33-45) */
static inline uint32_t lcg(uint32_t* s) {
  *s = *s * 1103515245u + 12345u;
  return *s;
}

/* procgen.zig:55-70 */
static void place_tree(uint32_t* s, bm* w, uint64_t x, uint64_t y, uint64_t z) {
  uint64_t th = lcg(s) % 4u + 4u;
  bm_set(w, x + 1, y, z + 1, 15u | SOLID);
  for (uint64_t o = 0; o < th; o++) bm_set(w, x + 1, y + o, z + 1, (14u + lcg(s) % 3u) | SOLID);
  for (uint64_t a = 0; a < 3; a++)
    for (uint64_t b = 0; b < 3; b++)
      for (uint64_t c = 0; c < 3; c++)
        bm_set(w, x + a, y + th + b, z + c, (18u + lcg(s) % 2u) | SOLID);
}

/* procgen.zig:6-53 on an empty VoxelBrickmap(dim, 8): chunks[(dim/8)^3]
 * zeroed by the caller, pool room for pool_capacity_words / 512 blocks
 * (zeroed).  Returns 0, 1 (pool too small), 2 (a write outside the map) or
 * 3 (dim not a multiple of 8 in [16, 1024]). */
int or_procgen(uint32_t dim, float offset_x, float offset_y, uint32_t* chunks, uint32_t* pool,
               uint64_t pool_capacity_words, uint64_t* pool_words) {
  if (dim < 16 || dim > 1024 || dim % 8) return 3;
  bm w = {dim, dim / 8, chunks, pool, pool_capacity_words / 512u, 0, 0};
  uint32_t s = 0x46AE4Fu;
  for (uint64_t x = 0; x < dim; x++)
    for (uint64_t z = 0; z < dim; z++)
      for (uint64_t y = 0; y < 16; y++) bm_set(&w, x, y, z, 13u | SOLID);
  for (uint64_t x = 0; x < dim && !w.err; x++) {
    for (uint64_t z = 0; z < dim; z++) {
      uint64_t vh = or_procgen_height(dim, offset_x, offset_y, (uint32_t)x, (uint32_t)z);
      for (uint64_t h = 0; h < vh; h++) {
        bm_set(&w, x, h, z, (21u + lcg(&s) % 3u) | SOLID);
        if (h <= 15)
          bm_set(&w, x, h, z, (25u + lcg(&s) % 3u) | SOLID);
        else if (h == vh - 1 && h > 15)
          bm_set(&w, x, h, z, (lcg(&s) % 6u) | SOLID);
      }
      if (vh > 16) {
        if (bm_get(&w, x, vh, z) != 0) continue;
        if (lcg(&s) % 5u == 0) bm_set(&w, x, vh, z, 7u + lcg(&s) % 5u);
        if (lcg(&s) % 71u == 0) bm_set(&w, x, vh, z, (7u + 5u) | SOLID);
        if (lcg(&s) % 420u == 0 && x < 500 && z < 500 && x > 5 && z > 5)
          place_tree(&s, &w, x, vh, z);
      }
      (void)lcg(&s);
    }
  }
  if (pool_words) *pool_words = w.used * 512u;
  return w.err;
}
