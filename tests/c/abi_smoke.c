/* The C-ABI from C, as an engine's render shim would use libvt.so: context,
 * a dense world with one voxel, a render, the centre pick, an edit that
 * removes the voxel, error statuses; then the engine's world generator on an
 * empty brickmap world and a read-back of its chunk table.
 * Build: gcc -std=c11 -D__HIP_PLATFORM_AMD__ -I include -I /opt/rocm/include
 *        tests/c/abi_smoke.c -L <lib dir> -lvt -L /opt/rocm/lib -lamdhip64 -lm
 * Prints "ABI SMOKE OK" and exits 0 when every check holds. */
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vt.h"

#define CHECK(cond, ...)                                     \
  do {                                                       \
    if (!(cond)) {                                           \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);   \
      fprintf(stderr, __VA_ARGS__);                          \
      fprintf(stderr, " [last error: %s]\n", vt_last_error()); \
      return 1;                                              \
    }                                                        \
  } while (0)

/* Camera.UniformData bytes of an eye looking along the unit vector f:
 * columns right, up, forward, (0,0,0,1) -- the layout C_view is read in. */
static void look(vt_camera* c, const float e[3], const float f[3]) {
  float r[3] = {f[2], 0.0f, -f[0]}; /* cross((0,1,0), f) */
  float n = sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (n < 1e-6f) {
    r[0] = 1.0f;
    r[1] = 0.0f;
    r[2] = 0.0f;
    n = 1.0f;
  }
  for (int i = 0; i < 3; i++) r[i] /= n;
  const float u[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2],
                      f[0] * r[1] - f[1] * r[0]};
  memset(c, 0, sizeof *c);
  for (int i = 0; i < 3; i++) {
    c->position[i] = e[i];
    c->view[0 + i] = r[i];
    c->view[4 + i] = u[i];
    c->view[8 + i] = f[i];
  }
  c->position[3] = 1.0f;
  c->view[15] = 1.0f;
  c->fov = 1.2217305f; /* 70 degrees */
}

int main(void) {
  int st;
  vt_ctx* ctx = NULL;
  st = vt_create(&ctx, 0);
  CHECK(st == VT_OK && ctx, "vt_create: %s", vt_status_string(st));
  int cus = 0;
  char name[64];
  st = vt_device_info(ctx, &cus, name, sizeof name);
  CHECK(st == VT_OK && cus > 0, "vt_device_info");

  enum { D = 16 };
  static uint8_t grid[D * D * D];
  memset(grid, 0, sizeof grid);
  grid[5 + D * (3 + D * 7)] = 2; /* voxel (5, 3, 7), material 2 */
  uint32_t pal[256];
  for (int k = 0; k < 256; k++) pal[k] = 0xFF000000u | (uint32_t)(k * 0x010203u);
  vt_world* world = NULL;
  st = vt_world_create_dense_u8(ctx, grid, D, pal, &world);
  CHECK(st == VT_OK && world, "vt_world_create_dense_u8: %s", vt_status_string(st));

  /* looking straight down at the voxel */
  const float eye[3] = {5.5f, 12.0f, 7.5f};
  const float dir[3] = {0.0f, -0.99995f, 0.0099995f};
  vt_camera cam;
  look(&cam, eye, dir);

  enum { W = 8, H = 8 };
  vt_hit_record* d_hits = NULL;
  float* d_shaded = NULL;
  CHECK(hipMalloc((void**)&d_hits, W * H * sizeof(vt_hit_record)) == hipSuccess, "hipMalloc");
  CHECK(hipMalloc((void**)&d_shaded, W * H * 16) == hipSuccess, "hipMalloc");
  vt_render_params p;
  memset(&p, 0, sizeof p);
  p.width = W;
  p.height = H;
  p.max_steps_primary = 192;
  p.max_steps_shadow = 48;
  p.sun[0] = 0.4f;
  p.sun[1] = 0.8f;
  p.sun[2] = 0.3f;
  p.traversal = VT_TRAVERSAL_FLAT;
  vt_frame f;
  memset(&f, 0, sizeof f);
  f.shaded = d_shaded;
  f.hits = d_hits;
  st = vt_render(ctx, world, &cam, &p, &f, NULL);
  CHECK(st == VT_OK, "vt_render: %s", vt_status_string(st));
  st = vt_sync(ctx, NULL);
  CHECK(st == VT_OK, "vt_sync");
  vt_hit_record hits[W * H];
  CHECK(hipMemcpy(hits, d_hits, sizeof hits, hipMemcpyDeviceToHost) == hipSuccess, "copy");
  const uint32_t want = 5 + D * (3 + D * 7);
  const vt_hit_record* c = &hits[(H / 2) * W + W / 2];
  CHECK(c->voxel == want && (c->info & 7u) == 4u, "centre pixel: voxel %u face %u", c->voxel,
        c->info & 7u);
  CHECK(hits[0].voxel == 0xFFFFFFFFu, "corner pixel should see the sky");

  vt_pick_result pk;
  st = vt_pick(ctx, world, &cam, 64, &pk, NULL);
  CHECK(st == VT_OK && pk.data == pal[2] && pk.remove[0] == 5 && pk.remove[1] == 3 &&
            pk.remove[2] == 7 && pk.place[1] == 4,
        "vt_pick: data %08x remove %d %d %d", pk.data, pk.remove[0], pk.remove[1], pk.remove[2]);

  /* the terrain_edit remove: clear the picked cell, then the centre ray misses */
  const uint32_t xyz[3] = {(uint32_t)pk.remove[0], (uint32_t)pk.remove[1],
                           (uint32_t)pk.remove[2]};
  const uint32_t zero = 0;
  st = vt_world_set_voxels(ctx, world, xyz, &zero, 1, NULL);
  CHECK(st == VT_OK, "vt_world_set_voxels");
  uint32_t got = 7;
  st = vt_world_get_voxels(ctx, world, xyz, 1, &got, NULL);
  CHECK(st == VT_OK && got == 0, "vt_world_get_voxels: %u", got);
  st = vt_render(ctx, world, &cam, &p, &f, NULL);
  CHECK(st == VT_OK && vt_sync(ctx, NULL) == VT_OK, "vt_render after edit");
  CHECK(hipMemcpy(hits, d_hits, sizeof hits, hipMemcpyDeviceToHost) == hipSuccess, "copy");
  for (int i = 0; i < W * H; i++) CHECK(hits[i].voxel == 0xFFFFFFFFu, "pixel %d still hits", i);

  /* error statuses: no abort, a status and a message */
  st = vt_render(ctx, world, &cam, &p, NULL, NULL);
  CHECK(st == VT_ERR_NULL_POINTER && strlen(vt_last_error()) > 0, "null frame -> %d", st);
  p.width = 0;
  st = vt_render(ctx, world, &cam, &p, &f, NULL);
  CHECK(st == VT_ERR_INVALID_ARGUMENT, "zero width -> %d", st);
  vt_world* bad = (vt_world*)&st;
  st = vt_world_create_dense_u8(ctx, grid, 24, pal, &bad);
  CHECK(st == VT_ERR_INVALID_ARGUMENT && bad == NULL, "dim 24 -> %d", st);
  CHECK(strlen(vt_status_string(VT_ERR_DEVICE)) > 0, "status string");

  /* Game.init's world in the engine's layout: empty VoxelBrickmap, procgen */
  {
    const uint32_t D = 64, M = D / 8;
    static uint32_t zero_chunks[8 * 8 * 8];
    uint32_t* atlas = calloc(256u * 256u * 256u, 4);
    CHECK(atlas != NULL, "atlas");
    vt_world* bm = NULL;
    st = vt_world_create_brickmap(ctx, D, zero_chunks, NULL, 0, atlas, &bm);
    CHECK(st == VT_OK && bm, "vt_world_create_brickmap: %s", vt_status_string(st));
    st = vt_world_procgen(ctx, bm, 0.0f, 0.0f, NULL);
    CHECK(st == VT_OK, "vt_world_procgen: %s", vt_status_string(st));
    size_t words = 0;
    static uint32_t chunks[8 * 8 * 8];
    st = vt_world_read_brickmap(ctx, bm, chunks, NULL, 0, &words, NULL);
    /* the base layer allocates chunks (cx, 0, cz), (cx, 1, cz) first, in x-major order */
    CHECK(st == VT_OK && words >= 2u * M * M * 512u && chunks[0] == 1 && chunks[M] == 2 &&
              chunks[M * M] == 3,
          "vt_world_read_brickmap: %zu words, chunks %u %u %u", words, chunks[0], chunks[M],
          chunks[M * M]);
    const uint32_t base[3] = {3, 15, 40}, top[3] = {3, 16, 40};
    uint32_t v[2];
    st = vt_world_get_voxels(ctx, bm, base, 1, v, NULL);
    st |= vt_world_get_voxels(ctx, bm, top, 1, v + 1, NULL);
    /* y = 15: stone (13) or the sand-like 25..27; a 64^3 map's terrain stays below y = 16 */
    CHECK(st == VT_OK && v[0] >> 28 == 1u &&
              ((v[0] & 0xFFFFFFFu) == 13u || ((v[0] & 0xFFFFFFFu) >= 25u && (v[0] & 0xFFFFFFFu) <= 27u)) &&
              v[1] == 0u,
          "procgen voxels %08x %08x", v[0], v[1]);
    st = vt_world_procgen(ctx, bm, 0.0f, 0.0f, NULL);
    CHECK(st == VT_ERR_INVALID_ARGUMENT, "procgen on a non-empty world -> %d", st);
    CHECK(vt_world_destroy(bm) == VT_OK, "vt_world_destroy");
    free(atlas);
  }

  CHECK(vt_world_destroy(world) == VT_OK, "vt_world_destroy");
  CHECK(vt_destroy(ctx) == VT_OK, "vt_destroy");
  (void)hipFree(d_hits);
  (void)hipFree(d_shaded);
  printf("ABI SMOKE OK\n");
  return 0;
}
