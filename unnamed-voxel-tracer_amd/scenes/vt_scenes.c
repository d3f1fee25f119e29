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
#include <pthread.h>
#include <stdatomic.h>
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

/* Sprite shapes are independent of each other, so they are cut in parallel
 * (one solid-voxel bitmask per sprite); everything that consumes the random
 * streams and the stamping itself stay sequential, so the bytes do not depend
 * on the thread count. */
typedef struct sprite_job {
  const vts_spec* s;
  const uint64_t* keys;
  uint8_t* masks; /* n_sprites * size^3 / 8 */
  size_t q;       /* rank of the threshold value */
  atomic_uint next;
} sprite_job;

static void* sprite_worker(void* arg) {
  sprite_job* J = (sprite_job*)arg;
  const uint32_t S = J->s->sprite_size;
  const size_t cnt = (size_t)S * S * S;
  double* val = (double*)malloc(cnt * sizeof(double));
  double* srt = (double*)malloc(cnt * sizeof(double));
  if (!val || !srt) {
    free(val);
    free(srt);
    return (void*)1;
  }
  for (;;) {
    uint32_t i = atomic_fetch_add(&J->next, 1u);
    if (i >= J->s->n_sprites) break;
    uint64_t key = J->keys[i];
    for (uint32_t z = 0; z < S; z++)
      for (uint32_t y = 0; y < S; y++)
        for (uint32_t x = 0; x < S; x++) {
          double v = vnoise3((x + 0.5) * 0.25, (y + 0.5) * 0.25, (z + 0.5) * 0.25, key);
          v += 0.5 * vnoise3((x + 0.5) * 0.5, (y + 0.5) * 0.5, (z + 0.5) * 0.5, key ^ 0x55ull);
          val[x + S * (y + S * z)] = v;
        }
    memcpy(srt, val, cnt * sizeof(double));
    qsort(srt, cnt, sizeof(double), cmp_double);
    double thr = srt[J->q];
    uint8_t* m = J->masks + (size_t)i * (cnt / 8);
    memset(m, 0, cnt / 8);
    for (size_t k = 0; k < cnt; k++)
      if (val[k] > thr) m[k >> 3] |= (uint8_t)(1u << (k & 7));
  }
  free(val);
  free(srt);
  return NULL;
}

typedef struct terrain_job {
  const vts_spec* s;
  uint8_t* grid;
  atomic_uint next;
} terrain_job;

static void* terrain_worker(void* arg) {
  terrain_job* J = (terrain_job*)arg;
  const uint64_t n = J->s->dim;
  for (;;) {
    uint32_t z = atomic_fetch_add(&J->next, 1u);
    if (z >= n) break;
    for (uint64_t x = 0; x < n; x++) {
      uint32_t h = vts_height(J->s, (uint32_t)x, z);
      for (uint64_t y = 0; y < h; y++) {
        uint8_t m = (y + 1 == h) ? 1 : (y + 4 >= h ? 2 : 3);
        J->grid[x + n * (y + n * (uint64_t)z)] = m;
      }
    }
  }
  return NULL;
}

static int run_threads(void* (*fn)(void*), void* arg, int threads) {
  pthread_t th[64];
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  int started = 0, rc = 0;
  for (; started < threads; started++)
    if (pthread_create(&th[started], NULL, fn, arg) != 0) break;
  if (started == 0) return fn(arg) != NULL;
  for (int i = 0; i < started; i++) {
    void* r = NULL;
    pthread_join(th[i], &r);
    if (r) rc = 1;
  }
  return rc;
}

/* Fills grid (dim^3 bytes, caller-allocated) using up to `threads` threads;
 * the result does not depend on the thread count.  Returns 0 on success. */
int vts_generate_mt(const vts_spec* s, uint8_t* grid, int threads) {
  const uint64_t n = s->dim;
  if (n == 0 || !grid) return 1;
  memset(grid, 0, n * n * n);
  /* terrain */
  terrain_job tj;
  tj.s = s;
  tj.grid = grid;
  atomic_init(&tj.next, 0u);
  if (run_threads(terrain_worker, &tj, threads)) return 2;
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
    if (S == 0 || S > n || cnt % 8 != 0) return 3;
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * s->n_sprites);
    uint32_t* pos = (uint32_t*)malloc(sizeof(uint32_t) * 3 * s->n_sprites);
    uint8_t* mat = (uint8_t*)malloc(s->n_sprites);
    uint8_t* masks = (uint8_t*)malloc((cnt / 8) * s->n_sprites);
    if (!keys || !pos || !mat || !masks) {
      free(keys);
      free(pos);
      free(mat);
      free(masks);
      return 2;
    }
    rng rs = rng_stream(s->seed, 2);
    for (uint32_t i = 0; i < s->n_sprites; i++) {
      keys[i] = rng_next(&rs);
      pos[3 * i + 0] = rng_range(&rs, 0, (uint32_t)n - S);
      pos[3 * i + 2] = rng_range(&rs, 0, (uint32_t)n - S);
      mat[i] = (uint8_t)rng_range(&rs, 4, 8);
      pos[3 * i + 1] = vts_height(s, pos[3 * i] + S / 2, pos[3 * i + 2] + S / 2);
    }
    sprite_job sj;
    sj.s = s;
    sj.keys = keys;
    sj.masks = masks;
    sj.q = (size_t)floor((1.0 - s->sprite_fill) * (double)cnt);
    if (sj.q >= cnt) sj.q = cnt - 1;
    atomic_init(&sj.next, 0u);
    int rc = run_threads(sprite_worker, &sj, threads);
    for (uint32_t i = 0; rc == 0 && i < s->n_sprites; i++) {
      const uint32_t sx = pos[3 * i], sy = pos[3 * i + 1], sz = pos[3 * i + 2];
      const uint8_t* m = masks + (size_t)i * (cnt / 8);
      for (uint32_t z = 0; z < S; z++)
        for (uint32_t y = 0; y < S; y++)
          for (uint32_t x = 0; x < S; x++) {
            const size_t k = x + S * (y + (size_t)S * z);
            uint64_t gy = (uint64_t)sy + y;
            if (gy >= n) continue;
            if ((m[k >> 3] >> (k & 7)) & 1u)
              grid[(sx + x) + n * (gy + n * (uint64_t)(sz + z))] = mat[i];
          }
    }
    free(keys);
    free(pos);
    free(mat);
    free(masks);
    if (rc) return 2;
  }
  return 0;
}

int vts_generate(const vts_spec* s, uint8_t* grid) { return vts_generate_mt(s, grid, 1); }
