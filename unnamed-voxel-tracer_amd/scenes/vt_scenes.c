/*
 * vt_scenes.c -- seeded synthetic voxel worlds for the benchmark workloads.
 *
 * Dense uint8 grids, grid[x + dim*(y + dim*z)], 0 = empty, 1..8 materials:
 *   terrain   heightmap from 4-octave 2D value-noise fBm (lacunarity 2,
 *             gain 0.5, quintic interpolation), height = floor(h_base +
 *             h_amp * n); top layer material 1, next 3 layers material 2,
 *             everything below material 3;
 *   blocks    solid cubes, edge uniform in [block_min, block_max], position
 *             uniform inside the grid, material uniform in 4..8, overwriting;
 *   sprites   size^3 models cut from 3D value noise at the (1-fill)
 *             quantile of each model, placed at uniform (x,z) on the terrain
 *             surface, material uniform in 4..8, solid voxels overwrite.
 * Every random stream is splitmix64 seeded from (seed, stream id), and all
 * noise is evaluated in binary64 with separately rounded operations, so a
 * spec always produces the same bytes on any x86-64 host.
 *
 * The reference generates its world with FastNoiseLite fBm in
 * src/procgen.zig:6-53; these workloads replace that generator with a fully
 * specified one so grids can be pinned by hash.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct vts_spec {
  uint32_t dim;
  uint64_t seed;
  double xz_scale;  /* noise lattice coordinates = (x, z) * base_freq / xz_scale */
  double base_freq; /* lattice cells per voxel at octave 0                        */
  double h_base, h_amp;
  uint32_t n_blocks, block_min, block_max;
  uint32_t n_sprites, sprite_size;
  double sprite_fill; /* fraction of a sprite's voxels that are solid          */
} vts_spec;

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

typedef struct rng {
  uint64_t s;
} rng;

static inline uint64_t rng_next(rng* r) {
  r->s += 0x9E3779B97F4A7C15ull;
  return mix64(r->s);
}
static inline rng rng_stream(uint64_t seed, uint64_t stream) {
  rng r = {mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ull))};
  return r;
}
/* uniform integer in [lo, hi] */
static inline uint32_t rng_range(rng* r, uint32_t lo, uint32_t hi) {
  return lo + (uint32_t)(rng_next(r) % (uint64_t)(hi - lo + 1u));
}

/* lattice value in [0,1) */
static inline double lattice(int64_t x, int64_t y, int64_t z, uint64_t key) {
  uint64_t h = mix64((uint64_t)x * 0x8CB92BA72F3D8DD7ull ^ (uint64_t)y * 0xD6E8FEB86659FD93ull ^
                     (uint64_t)z * 0xA0761D6478BD642Full ^ key);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

static inline double quintic(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
static inline double lerp(double a, double b, double t) { return a + (b - a) * t; }

static double vnoise2(double x, double z, uint64_t key) {
  double fx0 = floor(x), fz0 = floor(z);
  int64_t x0 = (int64_t)fx0, z0 = (int64_t)fz0;
  double u = quintic(x - fx0), v = quintic(z - fz0);
  double a = lattice(x0, 0, z0, key), b = lattice(x0 + 1, 0, z0, key);
  double c = lattice(x0, 0, z0 + 1, key), d = lattice(x0 + 1, 0, z0 + 1, key);
  return lerp(lerp(a, b, u), lerp(c, d, u), v);
}

static double vnoise3(double x, double y, double z, uint64_t key) {
  double fx0 = floor(x), fy0 = floor(y), fz0 = floor(z);
  int64_t x0 = (int64_t)fx0, y0 = (int64_t)fy0, z0 = (int64_t)fz0;
  double u = quintic(x - fx0), v = quintic(y - fy0), w = quintic(z - fz0);
  double c000 = lattice(x0, y0, z0, key), c100 = lattice(x0 + 1, y0, z0, key);
  double c010 = lattice(x0, y0 + 1, z0, key), c110 = lattice(x0 + 1, y0 + 1, z0, key);
  double c001 = lattice(x0, y0, z0 + 1, key), c101 = lattice(x0 + 1, y0, z0 + 1, key);
  double c011 = lattice(x0, y0 + 1, z0 + 1, key), c111 = lattice(x0 + 1, y0 + 1, z0 + 1, key);
  double x00 = lerp(c000, c100, u), x10 = lerp(c010, c110, u);
  double x01 = lerp(c001, c101, u), x11 = lerp(c011, c111, u);
  return lerp(lerp(x00, x10, v), lerp(x01, x11, v), w);
}

/* 4-octave fBm normalised to [0,1) */
double vts_fbm2(const vts_spec* s, double x, double z) {
  uint64_t key = mix64(s->seed ^ 0x7E77A1ull);
  double f = s->base_freq / s->xz_scale, amp = 1.0, sum = 0.0, norm = 0.0;
  for (int o = 0; o < 4; o++) {
    sum += amp * vnoise2(x * f, z * f, key + (uint64_t)o * 0x1000193ull);
    norm += amp;
    amp *= 0.5;
    f *= 2.0;
  }
  return sum / norm;
}

/* terrain height (first empty y) of column (x, z), clamped to [0, dim] */
uint32_t vts_height(const vts_spec* s, uint32_t x, uint32_t z) {
  double h = floor(s->h_base + s->h_amp * vts_fbm2(s, (double)x, (double)z));
  if (h < 0.0) h = 0.0;
  if (h > (double)s->dim) h = (double)s->dim;
  return (uint32_t)h;
}

static int cmp_double(const void* a, const void* b) {
  double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

/* Fills grid (dim^3 bytes, caller-allocated).  Returns 0 on success. */
int vts_generate(const vts_spec* s, uint8_t* grid) {
  const uint64_t n = s->dim;
  if (n == 0 || !grid) return 1;
  memset(grid, 0, n * n * n);
  /* terrain */
  for (uint64_t z = 0; z < n; z++) {
    for (uint64_t x = 0; x < n; x++) {
      uint32_t h = vts_height(s, (uint32_t)x, (uint32_t)z);
      for (uint64_t y = 0; y < h; y++) {
        uint8_t m = (y + 1 == h) ? 1 : (y + 4 >= h ? 2 : 3);
        grid[x + n * (y + n * z)] = m;
      }
    }
  }
  /* solid blocks */
  rng rb = rng_stream(s->seed, 1);
  for (uint32_t i = 0; i < s->n_blocks; i++) {
    uint32_t e = rng_range(&rb, s->block_min, s->block_max);
    if (e > n) e = (uint32_t)n;
    uint32_t bx = rng_range(&rb, 0, (uint32_t)n - e);
    uint32_t by = rng_range(&rb, 0, (uint32_t)n - e);
    uint32_t bz = rng_range(&rb, 0, (uint32_t)n - e);
    uint8_t m = (uint8_t)rng_range(&rb, 4, 8);
    for (uint64_t z = bz; z < bz + e; z++)
      for (uint64_t y = by; y < by + e; y++)
        for (uint64_t x = bx; x < bx + e; x++) grid[x + n * (y + n * z)] = m;
  }
  /* sprites */
  if (s->n_sprites) {
    const uint32_t S = s->sprite_size;
    const size_t cnt = (size_t)S * S * S;
    double* val = (double*)malloc(cnt * sizeof(double));
    double* srt = (double*)malloc(cnt * sizeof(double));
    if (!val || !srt) {
      free(val);
      free(srt);
      return 2;
    }
    rng rs = rng_stream(s->seed, 2);
    size_t q = (size_t)floor((1.0 - s->sprite_fill) * (double)cnt);
    if (q >= cnt) q = cnt - 1;
    for (uint32_t i = 0; i < s->n_sprites; i++) {
      uint64_t key = rng_next(&rs);
      uint32_t sx = rng_range(&rs, 0, (uint32_t)n - S);
      uint32_t sz = rng_range(&rs, 0, (uint32_t)n - S);
      uint8_t m = (uint8_t)rng_range(&rs, 4, 8);
      uint32_t sy = vts_height(s, sx + S / 2, sz + S / 2);
      for (uint32_t z = 0; z < S; z++)
        for (uint32_t y = 0; y < S; y++)
          for (uint32_t x = 0; x < S; x++) {
            double v = vnoise3((x + 0.5) * 0.25, (y + 0.5) * 0.25, (z + 0.5) * 0.25, key);
            v += 0.5 * vnoise3((x + 0.5) * 0.5, (y + 0.5) * 0.5, (z + 0.5) * 0.5, key ^ 0x55ull);
            val[x + S * (y + S * z)] = v;
          }
      memcpy(srt, val, cnt * sizeof(double));
      qsort(srt, cnt, sizeof(double), cmp_double);
      double thr = srt[q];
      for (uint32_t z = 0; z < S; z++)
        for (uint32_t y = 0; y < S; y++)
          for (uint32_t x = 0; x < S; x++) {
            uint64_t gy = (uint64_t)sy + y;
            if (gy >= n) continue;
            if (val[x + S * (y + S * z)] > thr)
              grid[(sx + x) + n * (gy + n * (uint64_t)(sz + z))] = m;
          }
    }
    free(val);
    free(srt);
  }
  return 0;
}
