/*
 * vt_oracle.c -- CPU restatement of the reference voxel ray-traversal path.
 *
 * TEST INFRASTRUCTURE ONLY (the test oracle and the timed CPU baseline); the
 * product library never links or calls it.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off; x86-64 SSE arithmetic, so every binary32
 * operation rounds to nearest-even exactly once).
 *
 * Each function cites the reference lines it follows (paths relative to the
 * reference repository root).  How each part is pinned (closed forms,
 * brute force, an independent geometric DDA) is listed in the header
 * comment of tests/test_oracle_pinning.py.
 */
#include "vt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* binary32 helpers                                                    */
/* ------------------------------------------------------------------ */

/* ivec3(float): truncate toward zero; NaN -> 0; saturating. */
static inline int32_t f2i(float x) {
  if (x != x) return 0;
  if (x >= 2147483648.0f) return INT32_MAX;
  if (x <= -2147483648.0f) return INT32_MIN;
  return (int32_t)x;
}

/* GLSL min/max: the non-NaN operand wins (IEEE minNum/maxNum). */
static inline float fmin_(float a, float b) { return fminf(a, b); }
static inline float fmax_(float a, float b) { return fmaxf(a, b); }

/* length(v) = sqrt(dot(v,v)), dot summed left to right. */
static inline float len3(float x, float y, float z) {
  float s = x * x;
  s = s + y * y;
  s = s + z * z;
  return sqrtf(s);
}

/* rgba8 UNORM store then load: round(clamp(x,0,1)*255)/255. */
static inline float unorm8(float x) {
  float c = fmin_(fmax_(x, 0.0f), 1.0f);
  float b = floorf(c * 255.0f + 0.5f);
  return b / 255.0f;
}

float or_tan_half_fov(float fov) { return (float)tan((double)(fov / 2.0f)); }

/* primary.comp.glsl:31-39 */
void or_raygen(const or_camera* cam, float tan_half, uint32_t W, uint32_t H, uint32_t px,
               uint32_t py, float o[3], float d[3]) {
  /* vec2 rayUV = vec2(pixelCoords) / vec2(size) * 2.0 - 1.0; */
  float u = (float)px / (float)W;
  float v = (float)py / (float)H;
  u = u * 2.0f - 1.0f;
  v = v * 2.0f - 1.0f;
  /* rayUV.y *= float(size.y) / float(size.x); */
  v = v * ((float)H / (float)W);
  /* rayUV *= tan(fov / 2); */
  u = u * tan_half;
  v = v * tan_half;
  /* normalize(C_view * vec4(rayUV, 1.0, 1.0)).xyz -- 4-component normalize */
  const float* M = cam->view;
  float r[4];
  for (int i = 0; i < 4; i++) {
    float acc = M[0 * 4 + i] * u;
    acc = acc + M[1 * 4 + i] * v;
    acc = acc + M[2 * 4 + i] * 1.0f;
    acc = acc + M[3 * 4 + i] * 1.0f;
    r[i] = acc;
  }
  float s = r[0] * r[0];
  s = s + r[1] * r[1];
  s = s + r[2] * r[2];
  s = s + r[3] * r[3];
  float l = sqrtf(s);
  d[0] = r[0] / l;
  d[1] = r[1] / l;
  d[2] = r[2] / l;
  /* vec3 rayOrigin = C_position.xyz; */
  o[0] = cam->position[0];
  o[1] = cam->position[1];
  o[2] = cam->position[2];
}

/* map.glsl:21-29 */
void or_intersect_aabb(const float o[3], const float d[3], const float bmin[3],
                       const float bmax[3], float* t_near, float* t_far) {
  float t1[3], t2[3];
  for (int i = 0; i < 3; i++) {
    float tmin = (bmin[i] - o[i]) / d[i];
    float tmax = (bmax[i] - o[i]) / d[i];
    t1[i] = fmin_(tmin, tmax);
    t2[i] = fmax_(tmin, tmax);
  }
  *t_near = fmax_(fmax_(t1[0], t1[1]), t1[2]);
  *t_far = fmin_(fmin_(t2[0], t2[1]), t2[2]);
}

/* primary.comp.glsl:42-43: rayOrigin + rayDir * max(intersection.x, 0) - EPSILON */
void or_entry_origin(const float o[3], const float d[3], uint32_t dim, float out[3]) {
  const float bmin[3] = {0.0f, 0.0f, 0.0f};
  const float bmax[3] = {(float)dim, (float)dim, (float)dim};
  float tn, tf;
  or_intersect_aabb(o, d, bmin, bmax, &tn, &tf);
  float tc = fmax_(tn, 0.0f);
  for (int i = 0; i < 3; i++) out[i] = (o[i] + d[i] * tc) - 0.001f;
}

/* ------------------------------------------------------------------ */
/* world lookups                                                        */
/* ------------------------------------------------------------------ */

/* map_getVoxel(ivec3(pos) >> 3) (map.glsl:31-47) for both world kinds;
 * for dense worlds the block value is the material id. */
static inline uint32_t get_block(const or_world* w, int32_t bx, int32_t by, int32_t bz) {
  if (w->kind == OR_WORLD_DENSE) {
    if (bx < 0 || by < 0 || bz < 0 || bx >= (int32_t)w->dim || by >= (int32_t)w->dim ||
        bz >= (int32_t)w->dim)
      return 0;
    return w->grid[(uint64_t)bx + (uint64_t)w->dim * ((uint64_t)by + (uint64_t)w->dim * bz)];
  }
  /* map_getChunkFlags(pos >> 3) (map.glsl:31-36) */
  int32_t M = (int32_t)(w->dim / 8u);
  int32_t cx = bx >> 3, cy = by >> 3, cz = bz >> 3;
  if (cx < 0 || cy < 0 || cz < 0 || cx >= M || cy >= M || cz >= M) return 0;
  uint32_t idx = w->chunks[cx + M * (cy + cz * M)];
  if (idx == 0) return 0;
  /* data[(blk_idx-1)*512 + (x%8) + ((z%8)*8 + (y%8))*8] (map.glsl:42-43) */
  uint64_t o = (uint64_t)(idx - 1u) * 512u + (uint64_t)(bx % 8) +
               (uint64_t)(((bz % 8) * 8 + (by % 8)) * 8);
  if (o >= w->pool_words) return 0; /* guarded by world validation */
  return w->pool[o];
}

/* map_getSubVoxel (map.glsl:57-60); dense worlds: the solid atlas cube of
 * colour palette[material]. */
static inline uint32_t get_sub(const or_world* w, uint32_t block, int32_t sx, int32_t sy,
                               int32_t sz) {
  if (w->kind == OR_WORLD_DENSE) return w->palette[block & 255u];
  uint32_t id = block & 0xFFFFFFFu;
  int32_t ox = (int32_t)((id & 31u) * 8u);
  int32_t oy = (int32_t)(((id / 32u) & 31u) * 8u);
  int32_t oz = (int32_t)(((id / 1024u) & 31u) * 8u);
  uint32_t x = (uint32_t)(sx + ox), y = (uint32_t)(sy + oy), z = (uint32_t)(sz + oz);
  return w->atlas[x + 256u * (y + 256u * z)];
}

/* normals[] (map.glsl:72-79) */
static const float NORMALS[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                                    {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};

/* traceMap (map.glsl:83-168), REF arithmetic. */
int or_trace_map(const or_world* w, const float o[3], const float d_in[3], int max_steps,
                 or_hit* hit, int32_t* path, int path_cap) {
  float d[3] = {d_in[0], d_in[1], d_in[2]};
  /* map.glsl:85-90: add a delta to prevent grid aligned rays */
  for (int i = 0; i < 3; i++)
    if (d[i] == 0.0f) d[i] = 0.001f;

  const int32_t B = (int32_t)(8u * w->dim); /* map.glsl:92 */
  int32_t sgn[3], posv[3];
  float inv[3];
  for (int i = 0; i < 3; i++) {
    sgn[i] = d[i] > 0.0f ? 1 : (d[i] < 0.0f ? -1 : 0); /* ivec3(sign(rayDir)) */
    posv[i] = (1 + sgn[i]) >> 1;                        /* (1 + raySign) >> 1 */
    inv[i] = 1.0f / d[i];                                /* 1.0 / rayDir */
  }
  int m = 0;
  int32_t g[3];
  float wc[3];
  for (int i = 0; i < 3; i++) { /* map.glsl:101-102 */
    float s8 = o[i] * 8.0f;
    g[i] = f2i(s8);
    wc[i] = s8 - (float)g[i];
  }
  int s = 0;

  memset(hit, 0, sizeof(*hit));
  hit->tie_margin = INFINITY;
  hit->hit_pos[0] = hit->hit_pos[1] = hit->hit_pos[2] = -1.0f;

  int step;
  for (step = 0; step < max_steps; step++) {
    if (g[0] >= B || g[1] >= B || g[2] >= B || g[0] < 0 || g[1] < 0 || g[2] < 0) break;
    /* uvec3 pos = uvec3(gridsCoords) + uvec3(withinGridCoords); */
    int32_t p[3];
    for (int i = 0; i < 3; i++) p[i] = g[i] + f2i(wc[i]);
    if (path && step < path_cap) {
      path[3 * step + 0] = p[0];
      path[3 * step + 1] = p[1];
      path[3 * step + 2] = p[2];
    }
    int32_t b[3] = {p[0] >> 3, p[1] >> 3, p[2] >> 3};
    if (w->kind == OR_WORLD_DENSE && b[0] >= 0 && b[1] >= 0 && b[2] >= 0 &&
        b[0] < (int32_t)w->dim && b[1] < (int32_t)w->dim && b[2] < (int32_t)w->dim) {
      hit->lookups++;
    }
    uint32_t block = get_block(w, b[0], b[1], b[2]);
    if (block != 0) {
      uint32_t sub = get_sub(w, block, p[0] % 8, p[1] % 8, p[2] % 8);
      if (sub != 0) { /* map.glsl:118-127 */
        int face = m == 0 ? 2 - posv[0] : (m == 1 ? 4 - posv[1] : 6 - posv[2]);
        hit->data = sub;
        hit->face = face;
        for (int i = 0; i < 3; i++) {
          hit->hit_pos[i] = (float)g[i] + wc[i];
          hit->normal[i] = NORMALS[face - 1][i];
          hit->block[i] = b[i];
        }
        hit->sub = (uint32_t)((p[0] & 7) + 8 * (p[1] & 7) + 64 * (p[2] & 7));
        hit->steps = step + 1;
        return 1;
      }
      if (s != 0) { /* map.glsl:131-135: back to the sub-voxel level */
        for (int i = 0; i < 3; i++) {
          g[i] += f2i(wc[i]);
          wc[i] = wc[i] - floorf(wc[i]);
        }
        s = 0;
      }
    } else if (s != 3) { /* map.glsl:140-144: up to the block level */
      for (int i = 0; i < 3; i++) {
        int32_t lo = g[i] & 7;
        wc[i] = wc[i] + (float)lo;
        g[i] -= lo;
      }
      s = 3;
    }
    /* dda stepping, map.glsl:157-162 */
    float t[3];
    for (int i = 0; i < 3; i++) t[i] = ((float)(posv[i] << s) - wc[i]) * inv[i];
    m = t[0] < t[1] ? (t[0] < t[2] ? 0 : 2) : (t[1] < t[2] ? 1 : 2);
    {
      float a = t[0], bb = t[1], c = t[2], lo, mid;
      if (a > bb) { float x = a; a = bb; bb = x; }
      if (bb > c) { float x = bb; bb = c; c = x; }
      if (a > bb) { float x = a; a = bb; bb = x; }
      lo = a;
      mid = bb;
      float mg = mid - lo;
      if (mg < hit->tie_margin) hit->tie_margin = mg;
    }
    g[m] += sgn[m] * (1 << s);
    float tm = t[m];
    for (int i = 0; i < 3; i++) wc[i] = wc[i] + d[i] * tm;
    wc[m] = (float)((1 - posv[m]) << s) * 0.999f;
  }
  hit->steps = step;
  hit->block[0] = hit->block[1] = hit->block[2] = -1;
  return 0; /* HitInfo(0, vec3(-1.0), vec3(0)) */
}

/* traceEntities (map.glsl:172-201); the body after the return on line 201 is
 * unreachable in the reference and not restated. */
static const float ENTITY_POS[5][3] = {
    {256.f, 21.f, 256.f}, {251.f, 21.f, 259.f}, {253.f, 21.f, 256.f},
    {251.f, 21.f, 256.f}, {257.f, 21.f, 261.f}};

int or_trace_entities(const float o[3], const float d[3], float max_distance) {
  float prevD = INFINITY;
  int id = -1;
  for (int i = 0; i < 5; i++) {
    const float* p = ENTITY_POS[i];
    float dist = len3(o[0] - p[0], o[1] - p[1], o[2] - p[2]);
    if (dist >= max_distance) continue;
    float bmax[3] = {p[0] + 1.0f, p[1] + 1.0f, p[2] + 1.0f};
    float tn, tf;
    or_intersect_aabb(o, d, p, bmax, &tn, &tf);
    if (tf >= tn && prevD >= tf) {
      id = i;
      prevD = tf;
    }
  }
  if (id >= 0) {
    const float* p = ENTITY_POS[id];
    float bmax[3] = {p[0] + 1.0f, p[1] + 1.0f, p[2] + 1.0f};
    float tn, tf;
    or_intersect_aabb(o, d, p, bmax, &tn, &tf);
    if (tf >= tn) return 1;
  }
  return 0;
}

/* SkyDome2 (camera.glsl:11-19) */
void or_sky(const float sun[3], const float rd[3], float rgb[3]) {
  float ls = len3(sun[0], sun[1], sun[2]);
  float lr = len3(rd[0], rd[1], rd[2]);
  float sn[3] = {sun[0] / ls, sun[1] / ls, sun[2] / ls};
  float rn[3] = {rd[0] / lr, rd[1] / lr, rd[2] / lr};
  float dt = sn[0] * rn[0];
  dt = dt + sn[1] * rn[1];
  dt = dt + sn[2] * rn[2];
  float sv = fmin_(fmax_(dt, 0.0f), 1.2f);
  const float base[3] = {0.6f, 0.71f, 0.75f};
  const float k[3] = {1.0f, 0.5f, 1.0f};
  const float g8[3] = {0.4f * 1.0f, 0.4f * 0.6f, 0.4f * 0.1f};
  const float g3[3] = {0.2f, 0.08f, 0.04f};
  float p8 = powf(sv, 8.0f), p3 = powf(sv, 3.0f);
  float ry = rd[1] * 0.2f;
  for (int i = 0; i < 3; i++) {
    float c = (base[i] - ry * k[i]) + 0.075f;
    c = c + g8[i] * p8;
    c = c + g3[i] * p3;
    rgb[i] = c;
  }
}

/* ------------------------------------------------------------------ */
/* full frame                                                           */
/* ------------------------------------------------------------------ */

typedef struct job {
  const or_world* w;
  const or_camera* cam;
  const or_params* p;
  const or_frame* f;
  uint32_t y0, y1;
  float tan_half;
  atomic_uint next_row;
  pthread_mutex_t lock;
} job;


/* ------------------------------------------------------------------ */
/* present pass, blit.fragment.glsl:15-21,32-35                        */
/* ------------------------------------------------------------------ */

/* Stores into an rgba8 target: clamp to [0,1], round to nearest. */
static uint32_t pack_unorm8(const float c[4]) {
  uint32_t r = 0;
  for (int i = 0; i < 4; i++) {
    float k = fmin_(fmax_(c[i], 0.0f), 1.0f);
    r |= (uint32_t)floorf(k * 255.0f + 0.5f) << (8 * i);
  }
  return r;
}

/* texPos = the pixel centre of the full-screen quad (blit.vertex.glsl:5-15). */
static void present_px(float c[4], uint32_t x, uint32_t y, uint32_t W, uint32_t H) {
  float tp[2] = {((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H};
  float d0 = tp[0] - 0.5f, d1 = tp[1] - 0.5f;
  float d2 = d0 * d0;
  d2 = d2 + d1 * d1;
  if (sqrtf(d2) <= 0.002f) { /* :33 mix(color, vec4(1,1,1,0.4), 0.5) */
    const float w[4] = {1.0f, 1.0f, 1.0f, 0.4f};
    for (int i = 0; i < 4; i++) c[i] = c[i] * (1.0f - 0.5f) + w[i] * 0.5f;
  }
  /* vignetteEffect (:18-21): uv *= 1 - uv; pow(uv.x*uv.y*15, 0.6*0.3) */
  float ux = tp[0] * (1.0f - tp[0]), uy = tp[1] * (1.0f - tp[1]);
  float g = powf(ux * uy * 15.0f, 0.6f * 0.3f);
  for (int i = 0; i < 4; i++) c[i] = g * c[i];
}

void or_present(const float* composite, uint32_t W, uint32_t H, float* out, uint32_t* display) {
  for (uint32_t y = 0; y < H; y++)
    for (uint32_t x = 0; x < W; x++) {
      size_t i = (size_t)y * W + x;
      float c[4];
      memcpy(c, composite + 4 * i, sizeof c);
      present_px(c, x, y, W, H);
      if (out) memcpy(out + 4 * i, c, sizeof c);
      if (display) display[i] = pack_unorm8(c);
    }
}

static void render_pixel(job* J, uint32_t px, uint32_t py, uint64_t st[8]) {
  const or_world* w = J->w;
  const or_params* p = J->p;
  const or_frame* f = J->f;
  const int unorm = (p->flags & 1u) != 0;
  const int lambert = (p->flags & 2u) != 0;
  const size_t pix = (size_t)py * p->width + px;

  /* ---- primary pass (primary.comp.glsl:31-68) ---- */
  float o[3], d[3], e[3];
  or_raygen(J->cam, J->tan_half, p->width, p->height, px, py, o, d);
  or_entry_origin(o, d, w->dim, e);
  or_hit h;
  int hit = or_trace_map(w, e, d, p->max_steps_primary, &h, NULL, 0);
  const int bitmask = p->traversal == 1u || p->traversal == 2u; /* 4^3 or 8^3 bricks */
  st[0] += (uint64_t)h.steps;
  /* bitmask traversal: the material byte is read once at a primary hit (its
   * word fetches, stats 4 and 5, depend on the kernel's cache and are counted
   * by the kernel alone); flat traversal: one gather per step, = steps. */
  if (bitmask) st[6] += (uint64_t)hit;

  float color[4], normal[4], position[4];
  if (hit) {
    for (int c = 0; c < 4; c++) color[c] = (float)((h.data >> (8 * c)) & 255u) / 255.0f;
    normal[0] = h.normal[0];
    normal[1] = h.normal[1];
    normal[2] = h.normal[2];
    normal[3] = 1.0f;
    for (int i = 0; i < 3; i++) position[i] = ceilf(h.hit_pos[i]) / 8.0f;
    position[3] = 1.0f;
  } else {
    float rgb[3];
    or_sky(p->sun, d, rgb);
    color[0] = rgb[0];
    color[1] = rgb[1];
    color[2] = rgb[2];
    color[3] = 1.0f;
    normal[0] = normal[1] = normal[2] = normal[3] = 1.0f;
    position[0] = position[1] = position[2] = position[3] = -1.0f;
  }
  if (unorm) { /* rgba8 images 0 and 1 (gbuffer.zig:10-11) */
    for (int c = 0; c < 4; c++) {
      color[c] = unorm8(color[c]);
      normal[c] = unorm8(normal[c]);
    }
  }

  /* ---- shadow pass (secondary.comp.glsl:25-50) ---- */
  float illum[4] = {0.f, 0.f, 0.f, 0.f};
  int occluded = 0;
  int shadow_steps = 0;
  float margin = h.tie_margin;
  if (!(position[0] < 0.0f || position[1] < 0.0f || position[2] < 0.0f)) {
    float so[3];
    for (int i = 0; i < 3; i++) so[i] = position[i] + normal[i] * 0.001f;
    or_hit sh;
    int shit = or_trace_map(w, so, p->sun, p->max_steps_shadow, &sh, NULL, 0);
    float hp8[3] = {sh.hit_pos[0] / 8.0f, sh.hit_pos[1] / 8.0f, sh.hit_pos[2] / 8.0f};
    float maxd = len3(so[0] - hp8[0], so[1] - hp8[1], so[2] - hp8[2]);
    int ent = or_trace_entities(so, p->sun, maxd);
    occluded = ent || shit;
    shadow_steps = sh.steps;
    if (sh.tie_margin < margin) margin = sh.tie_margin;
    st[1] += (uint64_t)sh.steps;
    st[2] += 1;
    st[3] += (uint64_t)occluded;
    st[7] += (uint64_t)ent;
    illum[0] = p->sun[0];
    illum[1] = p->sun[1];
    illum[2] = p->sun[2];
    illum[3] = occluded ? -0.3f : 0.3f;
    if (unorm)
      for (int c = 0; c < 4; c++) illum[c] = unorm8(illum[c]);
  }

  /* ---- outputs ---- */
  if (f->color) memcpy(f->color + 4 * pix, color, sizeof color);
  if (f->normal) memcpy(f->normal + 4 * pix, normal, sizeof normal);
  if (f->position) memcpy(f->position + 4 * pix, position, sizeof position);
  if (f->illumination) memcpy(f->illumination + 4 * pix, illum, sizeof illum);
  if (f->steps) f->steps[pix] = (uint32_t)h.steps | ((uint32_t)shadow_steps << 16);
  if (f->tie_margin) f->tie_margin[pix] = margin;

  float t = INFINITY;
  if (hit) {
    float hp[3] = {h.hit_pos[0] / 8.0f, h.hit_pos[1] / 8.0f, h.hit_pos[2] / 8.0f};
    t = len3(hp[0] - o[0], hp[1] - o[1], hp[2] - o[2]);
  }
  if (f->hits) {
    uint32_t voxel = 0xFFFFFFFFu, info = 0;
    if (hit) {
      voxel = (uint32_t)h.block[0] +
              w->dim * ((uint32_t)h.block[1] + w->dim * (uint32_t)h.block[2]);
      uint32_t mat;
      if (w->kind == OR_WORLD_DENSE)
        mat = get_block(w, h.block[0], h.block[1], h.block[2]) & 255u;
      else
        mat = get_block(w, h.block[0], h.block[1], h.block[2]) & 255u;
      info = (uint32_t)h.face | ((uint32_t)occluded << 3) | (h.sub << 4) | (mat << 16);
    }
    uint32_t tb;
    memcpy(&tb, &t, 4);
    f->hits[3 * pix + 0] = voxel;
    f->hits[3 * pix + 1] = tb;
    f->hits[3 * pix + 2] = info;
  }
  if (f->shaded) {
    float out[4];
    if (lambert) {
      if (hit) {
        float ls = len3(p->sun[0], p->sun[1], p->sun[2]);
        float l[3] = {p->sun[0] / ls, p->sun[1] / ls, p->sun[2] / ls};
        float nd = h.normal[0] * l[0];
        nd = nd + h.normal[1] * l[1];
        nd = nd + h.normal[2] * l[2];
        float fct = p->ambient + fmax_(nd, 0.0f) * (occluded ? 0.0f : 1.0f);
        size_t v = (size_t)h.block[0] +
                   (size_t)w->dim * ((size_t)h.block[1] + (size_t)w->dim * (size_t)h.block[2]);
        for (int c = 0; c < 3; c++) out[c] = p->albedo[3 * v + c] * fct;
        out[3] = 1.0f;
      } else {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
      }
    } else if (hit) { /* blit.fragment.glsl:30-31 */
      float sky[3];
      or_sky(p->sun, illum, sky);
      for (int c = 0; c < 3; c++) out[c] = color[c] + illum[3] * sky[c];
      out[3] = color[3] + illum[3] * 1.0f;
    } else {
      memcpy(out, color, sizeof out);
    }
    if ((p->flags & 16u) && !lambert) {
      present_px(out, px, py, p->width, p->height);
      if (f->display) f->display[pix] = pack_unorm8(out);
    }
    memcpy(f->shaded + 4 * pix, out, sizeof out);
  }
}

static void* worker(void* arg) {
  job* J = (job*)arg;
  uint64_t st[8] = {0};
  for (;;) {
    uint32_t y = atomic_fetch_add(&J->next_row, 1u) + J->y0;
    if (y >= J->y1) break;
    for (uint32_t x = 0; x < J->p->width; x++) render_pixel(J, x, y, st);
  }
  if (J->f->stats) {
    pthread_mutex_lock(&J->lock);
    for (int i = 0; i < 8; i++) J->f->stats[i] += st[i];
    pthread_mutex_unlock(&J->lock);
  }
  return NULL;
}

void or_render(const or_world* w, const or_camera* cam, const or_params* p, const or_frame* f,
               uint32_t y0, uint32_t y1, int threads) {
  job J;
  J.w = w;
  J.cam = cam;
  J.p = p;
  J.f = f;
  J.y0 = y0;
  J.y1 = y1 > p->height ? p->height : y1;
  J.tan_half = or_tan_half_fov(cam->fov);
  atomic_init(&J.next_row, 0u);
  pthread_mutex_init(&J.lock, NULL);
  if (threads < 1) threads = 1;
  if (threads == 1) {
    worker(&J);
  } else {
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, worker, &J);
    for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&J.lock);
}

/* ------------------------------------------------------------------ */
/* Lambert x visibility shading and its FP64 gradient                   */
/* ------------------------------------------------------------------ */

void or_shade_lambert(const uint32_t* hits, uint32_t npix, const float* albedo,
                      const float sun[3], float ambient, float* rgba_out) {
  float ls = len3(sun[0], sun[1], sun[2]);
  float l[3] = {sun[0] / ls, sun[1] / ls, sun[2] / ls};
  for (uint32_t i = 0; i < npix; i++) {
    uint32_t voxel = hits[3 * i], info = hits[3 * i + 2];
    float* out = rgba_out + 4 * (size_t)i;
    if (voxel == 0xFFFFFFFFu) {
      out[0] = out[1] = out[2] = out[3] = 0.0f;
      continue;
    }
    const float* n = NORMALS[(info & 7u) - 1u];
    float nd = n[0] * l[0];
    nd = nd + n[1] * l[1];
    nd = nd + n[2] * l[2];
    float fct = ambient + fmax_(nd, 0.0f) * (((info >> 3) & 1u) ? 0.0f : 1.0f);
    for (int c = 0; c < 3; c++) out[c] = albedo[3 * (size_t)voxel + c] * fct;
    out[3] = 1.0f;
  }
}

double or_lambert_grad64(const uint32_t* hits, uint32_t npix, const float* albedo,
                         const float sun[3], float ambient, const float* target_rgba,
                         double* dalbedo, double* dsun) {
  double s[3] = {sun[0], sun[1], sun[2]};
  double ls = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
  double l[3] = {s[0] / ls, s[1] / ls, s[2] / ls};
  double loss = 0.0;
  for (uint32_t i = 0; i < npix; i++) {
    uint32_t voxel = hits[3 * i], info = hits[3 * i + 2];
    const float* T = target_rgba + 4 * (size_t)i;
    if (voxel == 0xFFFFFFFFu) {
      for (int c = 0; c < 3; c++) loss += (double)T[c] * (double)T[c];
      continue;
    }
    const float* n = NORMALS[(info & 7u) - 1u];
    double nd = n[0] * l[0] + n[1] * l[1] + n[2] * l[2];
    double vis = ((info >> 3) & 1u) ? 0.0 : 1.0;
    double fct = ambient + (nd > 0.0 ? nd : 0.0) * vis;
    double gsum = 0.0; /* sum_c dL/dI_c * A_c */
    for (int c = 0; c < 3; c++) {
      double A = albedo[3 * (size_t)voxel + c];
      double I = A * fct;
      double r = I - (double)T[c];
      loss += r * r;
      dalbedo[3 * (size_t)voxel + c] += 2.0 * r * fct;
      gsum += 2.0 * r * A;
    }
    if (nd > 0.0 && vis > 0.0) {
      /* d(n . s/|s|)/ds = (n - (n.l) l) / |s| */
      for (int k = 0; k < 3; k++) dsun[k] += gsum * (n[k] - nd * l[k]) / ls;
    }
  }
  return loss;
}

/* ------------------------------------------------------------------ */
/* centre-ray pick, terrain_edit.comp.glsl:11-24                        */
/* ------------------------------------------------------------------ */

void or_pick(const or_world* w, const or_camera* cam, int max_steps, or_pick_result* out) {
  /* rayDir = normalize(C_view * vec4(rayUV = 0, 1.0, 1.0)).xyz */
  const float* M = cam->view;
  float r[4];
  for (int i = 0; i < 4; i++) {
    float acc = M[0 * 4 + i] * 0.0f;
    acc = acc + M[1 * 4 + i] * 0.0f;
    acc = acc + M[2 * 4 + i] * 1.0f;
    acc = acc + M[3 * 4 + i] * 1.0f;
    r[i] = acc;
  }
  float q = r[0] * r[0];
  q = q + r[1] * r[1];
  q = q + r[2] * r[2];
  q = q + r[3] * r[3];
  float l = sqrtf(q);
  float d[3] = {r[0] / l, r[1] / l, r[2] / l};
  float o[3] = {cam->position[0], cam->position[1], cam->position[2]}, e[3];
  or_entry_origin(o, d, w->dim, e);
  or_hit h;
  memset(out, 0, sizeof *out);
  int hit = or_trace_map(w, e, d, max_steps, &h, NULL, 0);
  out->steps = (uint32_t)h.steps;
  if (hit) {
    out->data = h.data;
    out->face = (uint32_t)h.face;
    memcpy(out->hit_pos, h.hit_pos, sizeof out->hit_pos);
    memcpy(out->normal, h.normal, sizeof out->normal);
  } else { /* map.glsl:167 */
    out->hit_pos[0] = out->hit_pos[1] = out->hit_pos[2] = -1.0f;
  }
  for (int i = 0; i < 3; i++) { /* :21 ivec3(hit_pos + normal) >> 3, :23 ivec3(hit_pos) >> 3 */
    out->place[i] = f2i(out->hit_pos[i] + out->normal[i]) >> 3;
    out->remove[i] = f2i(out->hit_pos[i]) >> 3;
  }
}
