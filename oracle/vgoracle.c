/*
 * ORACLE — test infrastructure only (see vgoracle.h). Sequential binary32 restatement of the
 * reference's per-frame GPU cut. Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math,
 * SSE binary32 arithmetic with correctly rounded / and sqrt).
 *
 * Every function names the reference lines it restates. Arithmetic follows the reference kernels'
 * written order: mat * vec = ((c0*x + c1*y) + c2*z) + c3*w, length = sqrt((x*x + y*y) + z*z),
 * dot = (a.x*b.x + a.y*b.y) + a.z*b.z.
 */
#include "vgoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* GeometryPool.ts:96-140, constants.ts:2-20, lodBlend.ts:22-24 */
#define MESH_STRIDE 6
#define INSTANCE_FLOATS 20
#define CELL_FLOATS 8
#define MAX_LEVELS 48
#define CUT_COUNTERS 5
#define C_SELECTED 0
#define C_DRAWN 1
#define C_SURVIVED 2
#define C_WORK 3
#define C_REQUESTED 4
#define PER_MESH_COUNTERS 12
#define S_SCALE 0
#define S_FILL 2
#define S_RETRY 4
#define RETRY_ROUNDS 2
#define FAST_PATH_MAX_MESHLETS 4
#define LOD_FADE_STEPS 255u
#define FULL_FADE_BITS (255u << 16)
#define HZB_MAX_LEVELS 16
#define HZB_HEADER (HZB_MAX_LEVELS * 4 + 4)
#define HZB_LEVEL_COUNT_SLOT (HZB_MAX_LEVELS * 4)

/* Literals as the reference's WGSL sees them: JS doubles rounded to binary32. */
#define K_INF_HALF ((float)(1e30 * 0.5))
#define K_INF ((float)1e30)
#define K_MARGIN_LO ((float)(1 - 1e-3))
#define K_MARGIN_HI ((float)(1 + 1e-3))
#define K_STATE_FIXED 10000.0f

/*
 * OR_MUTANT (tests only; 0 in every build the tests check against): compiled with -DOR_MUTANT=k, the
 * oracle carries one plausible mistake, k-th of the list below, so a test can show that its known answers
 * catch it. 1-5: a capacity-controller constant (GeometryPool.ts:131-137, 1032-1069); 6-7: the prefix
 * (GeometryPool.ts:973-995); 8-11: the distance and size culls (GeometryPool.ts:742-757, 877-893).
 */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif
#define FILL_TARGET (OR_MUTANT == 1 ? 0.9f : 0.82f)
#define FILL_HIGH (OR_MUTANT == 2 ? 0.95f : 0.9f)
#define RETRY_MAX (OR_MUTANT == 3 ? 1e30f : 4.0f)
#define STEER_UP_MAX (OR_MUTANT == 4 ? 1.05f : 1.03f)
#define STEER_DOWN (OR_MUTANT == 5 ? 1.02f : 1.01f)
#define FILL_LOW 0.75f
#define RETRY_MIN 1.1f

typedef struct {
  float x, y, z, w;
} f4;

static inline f4 mat_mul(const float* m, f4 v) {
  f4 r;
  r.x = ((m[0] * v.x + m[4] * v.y) + m[8] * v.z) + m[12] * v.w;
  r.y = ((m[1] * v.x + m[5] * v.y) + m[9] * v.z) + m[13] * v.w;
  r.z = ((m[2] * v.x + m[6] * v.y) + m[10] * v.z) + m[14] * v.w;
  r.w = ((m[3] * v.x + m[7] * v.y) + m[11] * v.z) + m[15] * v.w;
  return r;
}

static inline float len3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
static inline float dot3(const float* n, float x, float y, float z) { return (n[0] * x + n[1] * y) + n[2] * z; }
static inline f4 point(const float* p) { f4 r = {p[0], p[1], p[2], 1.0f}; return r; }

/* fdlibm e_logf.c (Sun Microsystems, 1993): x = 2^k (1 + f), log(1 + f) = f - s(f - R(z)) with
 * s = f / (2 + f). Basic operations only, so every IEEE binary32 implementation gives the same bits. */
float or_logf(float x) {
  const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f, two25 = 3.355443200e+07f;
  const float Lg1 = 6.6666668653e-01f, Lg2 = 4.0000000596e-01f, Lg3 = 2.8571429849e-01f, Lg4 = 2.2222198546e-01f,
              Lg5 = 1.8183572590e-01f, Lg6 = 1.5313838422e-01f, Lg7 = 1.4798198640e-01f;
  uint32_t ix;
  memcpy(&ix, &x, 4);
  int32_t k = 0;
  if (ix >= 0x7f800000u) return x + x;            /* inf, nan, negative */
  if (ix < 0x00800000u) {                          /* zero or subnormal */
    if (ix == 0) return -1.0f / 0.0f;
    k -= 25;
    x *= two25;
    memcpy(&ix, &x, 4);
  }
  k += (int32_t)(ix >> 23) - 127;
  ix &= 0x007fffffu;
  uint32_t i = (ix + (0x95f64u << 3)) & 0x800000u;
  uint32_t mbits = ix | (i ^ 0x3f800000u);         /* normalise x or x/2 */
  float m;
  memcpy(&m, &mbits, 4);
  k += (int32_t)(i >> 23);
  float f = m - 1.0f;
  float dk = (float)k;
  float s = f / (2.0f + f);
  float z = s * s;
  float w = z * z;
  float t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  float t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  float R = t2 + t1;
  float hfsq = 0.5f * f * f;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

uint32_t or_counter_words(uint32_t slots) { return PER_MESH_COUNTERS + 2u * slots * 4u; }

/* ---------------------------------------------------------------- per-frame context */

typedef struct {
  const or_scene* sc;
  const or_view* v;
  uint32_t* state;
  uint32_t* counters;
  uint32_t* args;
  uint32_t* draw;
  uint32_t* cell_list;   /* 4 per visible cell */
  uint32_t* work[2];     /* 4 per item */
  uint32_t work_cap;
  float low[2], high[2], fade_scale;
  int fp64;
  or_stats* st;
  /* two-phase occlusion (the library's design, not the reference's): the cut records its camera
     occlusion rejects of the final round; the second phase (phase 2) re-tests them */
  int phase;
  int record;
  uint8_t* reject_inst;   /* per instance */
  uint32_t* reject_mesh;  /* 5 words per record: instance, meshlet, fades, mesh slot, scale bits */
  uint32_t reject_n, reject_cap, reject_overflow;
} ctx_t;

static int cell_has_reject(const ctx_t* c, uint32_t first, uint32_t count) {
  for (uint32_t i = first; i < first + count; i++)
    if (c->reject_inst[i]) return 1;
  return 0;
}

static void clear_rejects(ctx_t* c) {
  if (!c->record) return;
  memset(c->reject_inst, 0, c->sc->instance_count);
  c->reject_n = 0;
  c->reject_overflow = 0;
}

static inline uint32_t* counter(ctx_t* c, int cut, int k) { return &c->counters[2 + cut * CUT_COUNTERS + k]; }
static inline uint32_t* mesh_count(ctx_t* c, int cut, uint32_t slot, uint32_t part) {
  return &c->counters[PER_MESH_COUNTERS + cut * c->sc->slots * 4 + slot * 4 + part * 2];
}
static inline uint32_t* mesh_start(ctx_t* c, int cut, uint32_t slot, uint32_t part) {
  return &c->counters[PER_MESH_COUNTERS + cut * c->sc->slots * 4 + 1 + slot * 4 + part * 2];
}
static inline const float* mesh_record(const or_scene* s, uint32_t slot, int k) {
  return &s->mesh_table[(slot * MESH_STRIDE + k) * 4];
}

/* getState / setState, GeometryPool.ts:622-624 */
static inline float get_state(ctx_t* c, int i) { return (float)c->state[i] / K_STATE_FIXED; }
static inline void set_state(ctx_t* c, int i, float value) { c->state[i] = (uint32_t)rintf(value * K_STATE_FIXED); }

/* loadBands, GeometryPool.ts:652-665 (smoothstep as WGSL defines it) */
static void load_bands(ctx_t* c) {
  const or_view* v = c->v;
  float camera_scale = get_state(c, S_SCALE);
  float shadow_scale = get_state(c, S_SCALE + 1);
  float x = v->error_threshold * camera_scale;
  float t = fminf(fmaxf((x - 1.25f) / (2.5f - 1.25f), 0.0f), 1.0f);
  float smooth = (t * t) * (3.0f - 2.0f * t);
  float width = 1.0f + (v->lod_blend_max - 1.0f) * smooth;
  int blending = width > 1.001f;
  c->fade_scale = blending ? 255.0f / or_logf(width) : 0.0f;
  float t0 = v->error_threshold * camera_scale;
  float t1 = (v->error_threshold * v->shadow_error_scale) * fmaxf(camera_scale, shadow_scale);
  c->low[0] = t0;
  c->high[0] = blending ? t0 * width : t0;
  c->low[1] = t1;
  c->high[1] = blending ? t1 * width : t1;
}

/* modelOf, GeometryPool.ts:669-675 */
static inline float model_scale(const float* m) {
  return fmaxf(len3(m[0], m[1], m[2]), fmaxf(len3(m[4], m[5], m[6]), len3(m[8], m[9], m[10])));
}

/* frustumDistance, GeometryPool.ts:690-698. mode 0: plain, 1: swept (x 1), 2: swept x select(flag, 1, 0). */
static float frustum_distance(const or_view* v, float cx, float cy, float cz, int mode, int flag) {
  float best = 1e30f;
  for (int p = 0; p < 6; p++) {
    const float* n = &v->planes[p * 4];
    float d = dot3(n, cx, cy, cz) + n[3];
    if (mode == 1) d = d + fmaxf(dot3(n, v->sweep[0], v->sweep[1], v->sweep[2]), 0.0f) * 1.0f;
    if (mode == 2) d = d + fmaxf(dot3(n, v->sweep[0], v->sweep[1], v->sweep[2]), 0.0f) * (flag ? 1.0f : 0.0f);
    best = fminf(best, d);
  }
  return best;
}

/* projectedError, GeometryPool.ts:678-683 */
static inline float projected_error(const or_view* v, const float* m, float scale, const float* sphere, float error) {
  f4 vc = mat_mul(v->view, mat_mul(m, point(sphere)));
  float distance = fmaxf(len3(vc.x, vc.y, vc.z) - sphere[3] * scale, v->near_plane);
  float pixels = ((error * scale) / distance) * v->proj_scale;
  return error >= K_INF_HALF ? K_INF : pixels;
}

/* ---------------------------------------------------------------- HZB */

static uint32_t hzb_layout(uint32_t max_size, uint32_t* offsets, uint32_t* capacity) {
  uint32_t dim = max_size, total = HZB_HEADER, n = 0;
  for (;;) {
    if (offsets) offsets[n] = total;
    total += dim * dim;
    n++;
    if (dim == 1) break;
    dim = (dim + 1) / 2;
  }
  if (capacity) *capacity = n;
  return total;
}

uint32_t or_hzb_size(uint32_t max_size) { return hzb_layout(max_size, NULL, NULL); }

/* OcclusionCulling.ts:92-113 (layout), 284-300 (header), 302-350 (copy + max reduction) */
int or_hzb_build(const float* depth, uint32_t width, uint32_t height, uint32_t max_size, float* hzb) {
  if (width == 0 || height == 0 || width > max_size || height > max_size) return -1;
  uint32_t offsets[HZB_MAX_LEVELS], capacity;
  uint32_t total = hzb_layout(max_size, offsets, &capacity);
  memset(hzb, 0, sizeof(float) * total);
  uint32_t lw = width, lh = height, levels = 0, w[HZB_MAX_LEVELS], h[HZB_MAX_LEVELS];
  for (uint32_t i = 0; i < capacity; i++) {
    hzb[i * 4] = (float)offsets[i];
    hzb[i * 4 + 1] = (float)lw;
    hzb[i * 4 + 2] = (float)lh;
    w[i] = lw;
    h[i] = lh;
    levels++;
    if (lw == 1 && lh == 1) break;
    lw = lw > 1 ? (lw + 1) / 2 : 1;
    lh = lh > 1 ? (lh + 1) / 2 : 1;
  }
  hzb[HZB_LEVEL_COUNT_SLOT] = (float)levels;
  memcpy(&hzb[offsets[0]], depth, sizeof(float) * width * height);
  for (uint32_t i = 1; i < levels; i++) {
    const float* src = &hzb[offsets[i - 1]];
    float* dst = &hzb[offsets[i]];
    uint32_t sw = w[i - 1], sh = h[i - 1];
    for (uint32_t y = 0; y < h[i]; y++)
      for (uint32_t x = 0; x < w[i]; x++) {
        uint32_t x0 = x * 2, y0 = y * 2;
        uint32_t x1 = x0 + 1 < sw - 1 ? x0 + 1 : sw - 1;
        uint32_t y1 = y0 + 1 < sh - 1 ? y0 + 1 : sh - 1;
        float a = fmaxf(src[y0 * sw + x0], src[y0 * sw + x1]);
        float b = fmaxf(src[y1 * sw + x0], src[y1 * sw + x1]);
        dst[y * w[i] + x] = fmaxf(a, b);
      }
  }
  return 0;
}

static inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

/* ceil(log2(n)) for n >= 1 with integer operations. */
static inline uint32_t ceil_log2(uint32_t n) { return n <= 1 ? 0u : 32u - (uint32_t)__builtin_clz(n - 1u); }

/* occluded, OcclusionCulling.ts:119-170 */
int or_occluded(const float* P, float near_plane, const float* hzb, const float* vc, float radius) {
  float depth = -vc[2];
  float near_depth = depth - radius;
  float far_depth = depth + radius;
  if (near_depth <= near_plane) return 0; /* behind: kept, and nothing is sampled */
  float x_lo = vc[0] - radius, x_hi = vc[0] + radius, y_lo = vc[1] - radius, y_hi = vc[1] + radius;
  float sx = P[0], sy = P[5];
  float min_x = fminf(x_lo / near_depth, x_lo / far_depth) * sx - P[8];
  float max_x = fmaxf(x_hi / near_depth, x_hi / far_depth) * sx - P[8];
  float min_y = fminf(y_lo / near_depth, y_lo / far_depth) * sy - P[9];
  float max_y = fmaxf(y_hi / near_depth, y_hi / far_depth) * sy - P[9];
  float min_z = P[14] / near_depth - P[10];
  float W = hzb[1], H = hzb[2];
  float x0 = clampf(floorf((min_x * 0.5f + 0.5f) * W), 0.0f, W - 1.0f);
  float x1 = clampf(floorf((max_x * 0.5f + 0.5f) * W), 0.0f, W - 1.0f);
  float y0 = clampf(floorf((0.5f - max_y * 0.5f) * H), 0.0f, H - 1.0f);
  float y1 = clampf(floorf((0.5f - min_y * 0.5f) * H), 0.0f, H - 1.0f);
  uint32_t extent = (uint32_t)fmaxf(x1 - x0, y1 - y0) + 1u;
  uint32_t level = ceil_log2(extent);
  uint32_t top = (uint32_t)hzb[HZB_LEVEL_COUNT_SLOT] - 1u;
  if (level > top) level = top;
  float scale = (float)(1u << level);
  float offset = hzb[level * 4], lw = hzb[level * 4 + 1], lh = hzb[level * 4 + 2];
#define SAMPLE(fx, fy) \
  hzb[(uint32_t)(offset + clampf(floorf((fy) / scale), 0.0f, lh - 1.0f) * lw + clampf(floorf((fx) / scale), 0.0f, lw - 1.0f))]
  float m = fmaxf(fmaxf(SAMPLE(x0, y0), SAMPLE(x1, y0)), fmaxf(SAMPLE(x0, y1), SAMPLE(x1, y1)));
#undef SAMPLE
  return min_z < 1.0f && min_z > m + 1e-6f;
}

/* occludedSphere, GeometryPool.ts:701-705 */
static int occluded_sphere(ctx_t* c, int on, float cx, float cy, float cz, float radius) {
  if (!on || !c->v->occlusion_active || !c->v->hzb) return 0;
  f4 vc = mat_mul(c->v->view, (f4){cx, cy, cz, 1.0f});
  float margin = (fabsf(vc.z) * c->high[0]) / c->v->proj_scale;
  float center[3] = {vc.x, vc.y, vc.z};
  c->st->occlusion_tests++;
  return or_occluded(c->v->proj, c->v->near_plane, c->v->hzb, center, radius + margin);
}

/* ---------------------------------------------------------------- fp64 check (cut.ts:26-48 form) */

static double pe64(const double* mv, double s, const or_view* v, const float* sphere, float error) {
  if (error >= 1e30 * 0.5) return INFINITY;
  double cx = sphere[0], cy = sphere[1], cz = sphere[2];
  double x = mv[0] * cx + mv[4] * cy + mv[8] * cz + mv[12];
  double y = mv[1] * cx + mv[5] * cy + mv[9] * cz + mv[13];
  double z = mv[2] * cx + mv[6] * cy + mv[10] * cz + mv[14];
  double d = hypot(hypot(x, y), z) - (double)sphere[3] * s;
  if (d < v->near_plane) d = v->near_plane;
  return ((double)error * s * (double)v->proj_scale) / d;
}

static int near64(double e, double t) { return fabs(e - t) <= 1e-6 * fabs(t); }

static void check64(ctx_t* c, int cut, const float* m, uint32_t meshlet, int decided32) {
  double mv[16], s = 0;
  for (int col = 0; col < 4; col++)
    for (int row = 0; row < 4; row++) {
      double acc = 0;
      for (int k = 0; k < 4; k++) acc += (double)c->v->view[k * 4 + row] * (double)m[col * 4 + k];
      mv[col * 4 + row] = acc;
    }
  for (int col = 0; col < 3; col++) {
    double l = hypot(hypot((double)m[col * 4], (double)m[col * 4 + 1]), (double)m[col * 4 + 2]);
    if (l > s) s = l;
  }
  const float* b = &c->sc->bounds[(size_t)meshlet * 16];
  double own = pe64(mv, s, c->v, b, b[8]);
  double parent = pe64(mv, s, c->v, b + 4, b[9]);
  double low = c->low[cut], high = c->high[cut];
  int in64 = own <= high && parent > low;
  int flagged = near64(own, high) || near64(parent, low);
  if (c->fade_scale > 0 && in64) {
    /* Step boundaries sit at errors low * exp((k + 1/2) / fadeScale): an error within 1e-6 relative of
     * one moves the step value by at most 1e-6 * fadeScale. Values outside (-1, 256) are clamped. */
    double scale = c->fade_scale;
    double vlo = log(fmax(own, 1e-30) / low) * scale, vhi = isinf(parent) ? 1e9 : log(fmax(parent, 1e-30) / low) * scale;
    double lo = fmin(255.0, fmax(0.0, nearbyint(vlo))), hi = isinf(parent) ? 255.0 : fmin(255.0, fmax(0.0, nearbyint(vhi)));
    in64 = hi > lo;
    double tol = 1e-6 * scale;
    if (vlo > -1.0 && vlo < 256.0 && fabs(vlo - floor(vlo) - 0.5) <= tol) flagged = 1;
    if (!isinf(parent) && vhi > -1.0 && vhi < 256.0 && fabs(vhi - floor(vhi) - 0.5) <= tol) flagged = 1;
  }
  c->st->lod_checked++;
  if (flagged) c->st->near_threshold++;
  if (in64 != decided32) {
    if (flagged) c->st->mismatch_near++;
    else c->st->mismatch_hard++;
  }
}

/* ---------------------------------------------------------------- passes */

/* resetNode, GeometryPool.ts:714-725 (every per-mesh word up to the slot capacity) */
static void reset(ctx_t* c) { memset(c->counters, 0, sizeof(uint32_t) * or_counter_words(c->sc->slots)); }

/* cellNode, GeometryPool.ts:728-760 */
static void cells_pass(ctx_t* c) {
  const or_scene* s = c->sc;
  const or_view* v = c->v;
  for (uint32_t cell = 0; cell < s->cell_count; cell++) {
    const float* sphere = &s->cells[cell * CELL_FLOATS];
    const float* record = sphere + 4;
    uint32_t slot = (uint32_t)record[1];
    c->st->cells_tested++;
    const float* flags = mesh_record(s, slot, 5);
    if (flags[0] == 0.0f) continue;
    float radius = sphere[3];
    if (v->frustum_culling == 1 && frustum_distance(v, sphere[0], sphere[1], sphere[2], 2, flags[1] > 0.5f) < -radius)
      continue;
    const float* levers = mesh_record(s, slot, 3);
    float min_pixels = fmaxf(v->min_pixel_radius, levers[1]);
    float max_radius = record[0];
    f4 vc = mat_mul(v->view, point(sphere));
    float view_distance = len3(vc.x, vc.y, vc.z);
    float gap = view_distance - radius;
    if ((gap > max_radius || OR_MUTANT == 8) && (max_radius * v->proj_scale) / fmaxf(gap, 1e-6f) < min_pixels) continue;
    if ((OR_MUTANT == 9 ? view_distance : gap) > levers[0]) continue;
    if (c->phase == 2 && !cell_has_reject(c, (uint32_t)record[2], (uint32_t)record[3])) continue;
    uint32_t at = c->counters[0]++;
    uint32_t* e = &c->cell_list[at * 4];
    e[0] = slot;
    e[1] = (uint32_t)record[2];
    e[2] = (uint32_t)record[3];
    e[3] = cell;
  }
}

/* selectLevels, GeometryPool.ts:771-843 */
static void select_levels(ctx_t* c, int cut, uint32_t slot, uint32_t level_count, const float* cache, int occlusion_flag,
                          const float* cull_bounds, uint32_t instance, const float* m, float scale, float pmin, float pmax) {
  const or_scene* s = c->sc;
  float low = c->low[cut], high = c->high[cut];
  uint32_t first = 0xffffffffu, last = 0, possible_n = 0, guaranteed_n = 0;
  for (uint32_t i = 0; i < level_count; i++) {
    const float* l0 = &cache[i * 8];
    const float* l1 = l0 + 4;
    uint32_t level_first = (uint32_t)l0[0];
    uint32_t size = (uint32_t)l0[1];
    int possible = (l0[2] * pmax <= high) && (l0[3] >= K_INF_HALF || l0[3] * pmin > low);
    int sure = (l1[0] * pmin <= low * K_MARGIN_LO) && (l1[1] >= K_INF_HALF || l1[1] * pmax > high * K_MARGIN_HI);
    if (possible) {
      if (level_first < first) first = level_first;
      if (level_first + size > last) last = level_first + size;
      possible_n += size;
      if (sure) guaranteed_n += size;
    }
  }
  f4 cc = mat_mul(m, point(cull_bounds));
  float cull_radius = cull_bounds[3] * scale;
  int fast = possible_n > 0 && possible_n == last - first && guaranteed_n == possible_n &&
             possible_n <= FAST_PATH_MAX_MESHLETS &&
             (c->v->frustum_culling == 0 || frustum_distance(c->v, cc.x, cc.y, cc.z, cut == 1 ? 1 : 0, 0) >= cull_radius);
  if (!(last > first)) return;
  (*counter(c, cut, C_SURVIVED))++;
  if (fast) {
    uint32_t base = *counter(c, cut, C_SELECTED);
    *counter(c, cut, C_SELECTED) += possible_n;
    c->st->clusters_tested[cut] += possible_n;
    for (uint32_t k = 0; k < FAST_PATH_MAX_MESHLETS; k++) {
      if (!(k < possible_n && base + k < s->max_meshlets)) continue;
      uint32_t meshlet = first + k;
      uint32_t tris = s->info[meshlet * 4];
      uint32_t* mc = mesh_count(c, cut, slot, 0);
      uint32_t first_index = *mc;
      *mc += tris * 3;
      uint32_t* d = &c->draw[(size_t)(base + k + cut * s->max_meshlets) * 4];
      d[0] = instance;
      d[1] = meshlet;
      d[2] = first_index;
      d[3] = tris | FULL_FADE_BITS;
    }
    if (c->fp64)
      for (uint32_t k = 0; k < possible_n; k++) check64(c, cut, m, first + k, 1);
  } else {
    uint32_t item = (*counter(c, cut, C_WORK))++;
    uint32_t* w = &c->work[cut][(size_t)item * 4];
    w[0] = instance;
    w[1] = first;
    w[2] = last;
    w[3] = slot | (occlusion_flag ? 0x80000000u : 0u);
  }
}

/* instancePass, GeometryPool.ts:846-913 */
static void instance_pass(ctx_t* c, int retry) {
  const or_scene* s = c->sc;
  const or_view* v = c->v;
  float cache[MAX_LEVELS * 8];
  uint32_t visible = c->counters[0];
  int redo_camera = 1, redo_shadow = 1;
  if (retry) {
    redo_camera = get_state(c, S_RETRY) > 0.5f;
    redo_shadow = get_state(c, S_RETRY + 1) > 0.5f;
  }
  for (uint32_t k = 0; k < visible; k++) {
    const uint32_t* cell = &c->cell_list[k * 4];
    uint32_t slot = cell[0];
    const float* levers = mesh_record(s, slot, 3);
    uint32_t level_count = (uint32_t)levers[2];
    if (level_count > MAX_LEVELS) level_count = MAX_LEVELS;
    uint32_t level_base = (uint32_t)levers[3];
    memcpy(cache, &s->mesh_table[level_base * 4], sizeof(float) * level_count * 8);
    const float* object_sphere = mesh_record(s, slot, 0);
    const float* lod_bounds = mesh_record(s, slot, 1);
    const float* cull_bounds = mesh_record(s, slot, 2);
    const float* flags = mesh_record(s, slot, 5);
    int shadow_on = flags[1] > 0.5f, occlusion_on = flags[2] > 0.5f;
    float min_pixels = OR_MUTANT == 11 ? v->min_pixel_radius : fmaxf(v->min_pixel_radius, levers[1]);
    float max_draw_distance = levers[0];
    for (uint32_t i = cell[1]; i < cell[1] + cell[2]; i++) {
      if (c->phase == 2 && !c->reject_inst[i]) continue;
      const float* m = &s->instances[(size_t)i * INSTANCE_FLOATS];
      float scale = model_scale(m);
      f4 center = mat_mul(m, point(object_sphere));
      float radius = object_sphere[3] * scale;
      int no_frustum = v->frustum_culling == 0;
      int in_camera = no_frustum || frustum_distance(v, center.x, center.y, center.z, 0, 0) >= -radius;
      int in_shadow = c->phase != 2 && shadow_on && (no_frustum || frustum_distance(v, center.x, center.y, center.z, 1, 0) >= -radius);
      if (retry) {
        in_camera = in_camera && redo_camera;
        in_shadow = in_shadow && redo_shadow;
      }
      c->st->instances_tested++;
      f4 vc = mat_mul(v->view, (f4){center.x, center.y, center.z, 1.0f});
      float view_distance = len3(vc.x, vc.y, vc.z);
      int too_small = view_distance > radius && (radius * v->proj_scale) / view_distance < min_pixels;
      int too_far = (OR_MUTANT == 10 ? view_distance : view_distance - radius) > max_draw_distance;
      if (!((in_camera || in_shadow) && !too_small && !too_far)) continue;
      if (in_camera && occluded_sphere(c, occlusion_on, center.x, center.y, center.z, radius)) {
        in_camera = 0;
        c->counters[1]++;
        if (c->record) c->reject_inst[i] = 1;
      }
      f4 lod_center = mat_mul(v->view, mat_mul(m, point(lod_bounds)));
      float d = len3(lod_center.x, lod_center.y, lod_center.z);
      float r = lod_bounds[3] * scale;
      float pmin = (scale * v->proj_scale) / fmaxf(d - r, v->near_plane);
      float pmax = (scale * v->proj_scale) / fmaxf(d + r, v->near_plane);
      if (in_camera) select_levels(c, 0, slot, level_count, cache, occlusion_on, cull_bounds, i, m, scale, pmin, pmax);
      if (in_shadow) select_levels(c, 1, slot, level_count, cache, 0, cull_bounds, i, m, scale, pmin, pmax);
    }
  }
}

/* step of the blend range, GeometryPool.ts:953 */
static inline uint32_t fade_step(float error, float low, float fade_scale) {
  float v = rintf(or_logf(fmaxf(error, 1e-30f) / low) * fade_scale);
  return (uint32_t)clampf(v, 0.0f, (float)LOD_FADE_STEPS);
}

/* meshletPass(cut).passNode, GeometryPool.ts:926-968 */
static void meshlet_pass(ctx_t* c, int cut) {
  const or_scene* s = c->sc;
  const or_view* v = c->v;
  uint32_t items = *counter(c, cut, C_WORK);
  float low = c->low[cut], high = c->high[cut], fade = c->fade_scale;
  for (uint32_t item = 0; item < items; item++) {
    const uint32_t* w = &c->work[cut][(size_t)item * 4];
    uint32_t instance = w[0], slot = w[3] & 0x7fffffffu;
    int occlusion_on = (w[3] >> 31) == 1u;
    const float* m = &s->instances[(size_t)instance * INSTANCE_FLOATS];
    float scale = model_scale(m);
    c->st->work_items[cut]++;
    c->st->clusters_tested[cut] += w[2] - w[1];
    for (uint32_t meshlet = w[1]; meshlet < w[2]; meshlet++) {
      const float* b = &s->bounds[(size_t)meshlet * 16];
      const float* cull = b + 12;
      f4 center = mat_mul(m, point(cull));
      int visible = v->frustum_culling == 0 ||
                    frustum_distance(v, center.x, center.y, center.z, cut == 1 ? 1 : 0, 0) >= -(cull[3] * scale);
      float own = projected_error(v, m, scale, b, b[8]);
      float parent = projected_error(v, m, scale, b + 4, b[9]);
      uint32_t fade_low = fade > 0.0f ? fade_step(own, low, fade) : 0u;
      uint32_t fade_high = fade > 0.0f ? fade_step(parent, low, fade) : LOD_FADE_STEPS;
      int lod = own <= high && parent > low && fade_high > fade_low;
      /* drawn = visible && lod; the occlusion test (and its count) only matters for clusters in the cut */
      if (cut == 0 && lod && visible) {
        visible = !occluded_sphere(c, occlusion_on, center.x, center.y, center.z, cull[3] * scale);
        if (!visible && c->record) {
          if (c->reject_n < c->reject_cap) {
            uint32_t* r = &c->reject_mesh[(size_t)c->reject_n * 5];
            r[0] = instance;
            r[1] = meshlet;
            r[2] = fade_low | (fade_high << 8);
            r[3] = slot;
            memcpy(&r[4], &scale, 4);
          } else {
            c->reject_overflow = 1;
          }
          c->reject_n++;
        }
      }
      if (c->fp64) check64(c, cut, m, meshlet, lod);
      if (!(visible && lod)) continue;
      uint32_t at = (*counter(c, cut, C_SELECTED))++;
      if (at >= s->max_meshlets) continue;
      uint32_t tris = s->info[meshlet * 4];
      uint32_t packed = tris | (fade_low << 8) | (fade_high << 16);
      uint32_t part = (packed & 0xffff00u) == FULL_FADE_BITS ? 0u : 1u;
      uint32_t* mc = mesh_count(c, cut, slot, part);
      uint32_t first_index = *mc;
      *mc += tris * 3;
      uint32_t* d = &c->draw[(size_t)(at + cut * s->max_meshlets) * 4];
      d[0] = instance;
      d[1] = meshlet;
      d[2] = first_index;
      d[3] = packed;
    }
  }
}

/* prefixPass, GeometryPool.ts:973-995 */
static void prefix_pass_from(ctx_t* c, int cut, uint32_t offset) {
  const or_scene* s = c->sc;
  uint32_t start = offset, capacity = s->cut_capacity[cut];
  for (uint32_t k = 0; k < s->mesh_slots * 2; k++) {
      /* slot-major: (slot 0, part 0), (slot 0, part 1), (slot 1, part 0), ... */
      const uint32_t i = OR_MUTANT == 7 ? k % s->mesh_slots : k / 2, part = OR_MUTANT == 7 ? k / s->mesh_slots : k % 2;
      uint32_t requested = *mesh_count(c, cut, i, part);
      *mesh_start(c, cut, i, part) = start;
      uint32_t room = start < capacity ? capacity - start : 0u;
      uint32_t* a = &c->args[cut * s->slots * 10 + part * 5 + i * 10];
      a[0] = requested < room || OR_MUTANT == 6 ? requested : room;
      a[1] = 1;
      a[2] = start;
      a[3] = 0;
      a[4] = 0;
      start += requested;
    }
  *counter(c, cut, C_REQUESTED) = start;
  uint32_t selected = *counter(c, cut, C_SELECTED);
  *counter(c, cut, C_DRAWN) = selected < s->max_meshlets ? selected : s->max_meshlets;
}

static void prefix_pass(ctx_t* c, int cut) { prefix_pass_from(c, cut, 0); }

/* fillOf, GeometryPool.ts:1027-1031 */
static float fill_of(ctx_t* c, int cut) {
  float requested = (float)*counter(c, cut, C_REQUESTED);
  float selected = (float)*counter(c, cut, C_SELECTED);
  if (*counter(c, cut, C_SELECTED) > c->sc->max_meshlets) c->st->draw_overflow[cut]++;
  return fmaxf(requested / (float)c->sc->cut_capacity[cut], selected / (float)c->sc->max_meshlets);
}

/* checkNode, GeometryPool.ts:1032-1054; returns whether any cut is redone */
static int check_pass(ctx_t* c) {
  int any = 0;
  for (int cut = 0; cut < 2; cut++) {
    float fill = fill_of(c, cut);
    float scale = get_state(c, S_SCALE + cut);
    set_state(c, S_RETRY + cut, 0.0f);
    if (fill > 1.0f && scale < c->v->lod_scale_limit) {
      set_state(c, S_SCALE + cut, fminf(scale * clampf(sqrtf(fill / FILL_TARGET), RETRY_MIN, RETRY_MAX), c->v->lod_scale_limit));
      set_state(c, S_RETRY + cut, 1.0f);
      any = 1;
      c->st->retries[cut]++;
      for (int k = 0; k < CUT_COUNTERS; k++) *counter(c, cut, k) = 0;
      if (cut == 0) c->counters[1] = 0;
      if (cut == 0) clear_rejects(c);
      for (uint32_t i = 0; i < c->sc->mesh_slots; i++) {
        *mesh_count(c, cut, i, 0) = 0;
        *mesh_count(c, cut, i, 1) = 0;
      }
    }
  }
  return any;
}

/* steerNode, GeometryPool.ts:1056-1069 */
static void steer_pass(ctx_t* c) {
  for (int cut = 0; cut < 2; cut++) {
    float fill = fill_of(c, cut);
    float scale = get_state(c, S_SCALE + cut);
    if (fill > FILL_HIGH) scale = scale * fminf(sqrtf(fill / FILL_TARGET), STEER_UP_MAX);
    else if (fill < FILL_LOW) scale = scale / STEER_DOWN;
    set_state(c, S_SCALE + cut, fmaxf(fminf(scale, c->v->lod_scale_limit), 1.0f));
    set_state(c, S_FILL + cut, fill);
  }
}

static int ctx_open(ctx_t* c, const or_scene* scene, const or_view* view, uint32_t* state, uint32_t* counters, uint32_t* draw_args,
                    uint32_t* draw_list, int fp64_check, or_stats* stats) {
  memset(c, 0, sizeof(*c));
  memset(stats, 0, sizeof(*stats));
  c->sc = scene;
  c->v = view;
  c->state = state;
  c->counters = counters;
  c->args = draw_args;
  c->draw = draw_list;
  c->fp64 = fp64_check;
  c->st = stats;
  c->cell_list = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (scene->cell_count + 1));
  c->work_cap = scene->instance_count + 1;
  c->work[0] = (uint32_t*)malloc(sizeof(uint32_t) * 4 * c->work_cap);
  c->work[1] = (uint32_t*)malloc(sizeof(uint32_t) * 4 * c->work_cap);
  if (!c->cell_list || !c->work[0] || !c->work[1]) return -1;
  memset(draw_args, 0, sizeof(uint32_t) * 2 * scene->slots * 2 * 5);
  return 0;
}

static void ctx_close(ctx_t* c) {
  free(c->cell_list);
  free(c->work[0]);
  free(c->work[1]);
}

/* GeometryPool.ts:1081-1092: reset, cells, instance, meshlets x2, prefix x2, 2 retry rounds, steer */
static void frame(ctx_t* c) {
  reset(c);
  load_bands(c);
  cells_pass(c);
  instance_pass(c, 0);
  meshlet_pass(c, 0);
  meshlet_pass(c, 1);
  prefix_pass(c, 0);
  prefix_pass(c, 1);
  for (int round = 0; round < RETRY_ROUNDS; round++) {
    int any = check_pass(c);
    load_bands(c);
    if (any) {
      instance_pass(c, 1);
      for (int cut = 0; cut < 2; cut++)
        if (get_state(c, S_RETRY + cut) > 0.5f) meshlet_pass(c, cut);
    }
    prefix_pass(c, 0);
    prefix_pass(c, 1);
  }
  steer_pass(c);
}

int or_cut(const or_scene* scene, const or_view* view, uint32_t* state, uint32_t* counters, uint32_t* draw_args,
           uint32_t* draw_list, int fp64_check, or_stats* stats) {
  ctx_t c;
  int rc = ctx_open(&c, scene, view, state, counters, draw_args, draw_list, fp64_check, stats);
  if (rc == 0) frame(&c);
  ctx_close(&c);
  return rc;
}

/*
 * Two-phase occlusion: the cut of `view1` (its HZB) recording the camera cut's occlusion rejects of the
 * final round, then the second phase against `view2`'s HZB (the same view otherwise): cells holding a
 * rejected instance, the instance pass on those instances (camera cut only), the meshlet pass on their
 * items, and the occlusion re-test of the rejected meshlets, at the cut's final band; appended to the
 * camera draw list after the cut's entries, regions after its requested indices.
 */
int or_cut_two_phase(const or_scene* scene, const or_view* view1, const or_view* view2, uint32_t* state, uint32_t* counters1,
                     uint32_t* args1, uint32_t* counters2, uint32_t* args2, uint32_t* draw_list, uint32_t reject_capacity,
                     or_stats* stats1, or_stats* stats2, uint32_t* rejects) {
  ctx_t c1, c2;
  int rc = ctx_open(&c1, scene, view1, state, counters1, args1, draw_list, 0, stats1);
  int rc2 = ctx_open(&c2, scene, view2, state, counters2, args2, draw_list, 0, stats2);
  c1.reject_inst = (uint8_t*)calloc(scene->instance_count + 1, 1);
  c1.reject_mesh = (uint32_t*)malloc(sizeof(uint32_t) * 5 * ((size_t)reject_capacity + 1));
  if (rc || rc2 || !c1.reject_inst || !c1.reject_mesh) {
    ctx_close(&c1);
    ctx_close(&c2);
    free(c1.reject_inst);
    free(c1.reject_mesh);
    return -1;
  }
  c1.record = 1;
  c1.reject_cap = reject_capacity;
  frame(&c1);
  /* the second phase */
  c2.phase = 2;
  c2.reject_inst = c1.reject_inst;
  reset(&c2);
  *counter(&c2, 0, C_SELECTED) = counters1[2 + C_SELECTED];
  c2.low[0] = c1.low[0];
  c2.high[0] = c1.high[0];
  c2.fade_scale = c1.fade_scale;
  cells_pass(&c2);
  instance_pass(&c2, 0);
  meshlet_pass(&c2, 0);
  const or_view* v = c2.v;
  uint32_t n = c1.reject_n < reject_capacity ? c1.reject_n : reject_capacity;
  for (uint32_t k = 0; k < n; k++) {
    const uint32_t* r = &c1.reject_mesh[(size_t)k * 5];
    uint32_t instance = r[0], meshlet = r[1], fades = r[2], slot = r[3];
    float scale;
    memcpy(&scale, &r[4], 4);
    const float* m = &scene->instances[(size_t)instance * INSTANCE_FLOATS];
    const float* cull = &scene->bounds[(size_t)meshlet * 16 + 12];
    f4 center = mat_mul(m, point(cull));
    (void)v;
    if (occluded_sphere(&c2, 1, center.x, center.y, center.z, cull[3] * scale)) continue;
    uint32_t at = (*counter(&c2, 0, C_SELECTED))++;
    if (at >= scene->max_meshlets) continue;
    uint32_t tris = scene->info[meshlet * 4];
    uint32_t packed = tris | (fades << 8);
    uint32_t part = (packed & 0xffff00u) == FULL_FADE_BITS ? 0u : 1u;
    uint32_t* mc = mesh_count(&c2, 0, slot, part);
    uint32_t first_index = *mc;
    *mc += tris * 3;
    uint32_t* d = &draw_list[(size_t)at * 4];
    d[0] = instance;
    d[1] = meshlet;
    d[2] = first_index;
    d[3] = packed;
  }
  prefix_pass_from(&c2, 0, counters1[2 + C_REQUESTED]);
  prefix_pass(&c2, 1);
  if (rejects) {
    uint32_t ni = 0;
    for (uint32_t i = 0; i < scene->instance_count; i++) ni += c1.reject_inst[i];
    rejects[0] = ni;
    rejects[1] = c1.reject_n;
    rejects[2] = c1.reject_overflow;
  }
  ctx_close(&c1);
  ctx_close(&c2);
  free(c1.reject_inst);
  free(c1.reject_mesh);
  return 0;
}

/* ---------------------------------------------------------------- triangle-budget histogram (our design) */

static uint32_t lower_bound_scaled(const float* thresholds, uint32_t bins, float scale, float x) {
  /* first b with thresholds[b] * scale >= x (bins when none) */
  uint32_t lo = 0, hi = bins;
  while (lo < hi) {
    uint32_t mid = (lo + hi) / 2;
    if (thresholds[mid] * scale >= x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

int or_budget_histogram(const or_scene* s, const or_view* v, const uint32_t* state_in, const float* thresholds,
                        uint32_t bins, uint64_t* triangles) {
  if (bins == 0) return -1;
  uint32_t state[8];
  memcpy(state, state_in, sizeof(state));
  uint64_t* diff = (uint64_t*)calloc(bins + 1, sizeof(uint64_t));
  uint32_t* counters = (uint32_t*)calloc(or_counter_words(s->slots), sizeof(uint32_t));
  ctx_t c;
  or_stats st;
  memset(&c, 0, sizeof(c));
  memset(&st, 0, sizeof(st));
  c.sc = s;
  c.st = &st;
  c.state = state;
  c.counters = counters;
  c.cell_list = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (s->cell_count + 1));
  /* The band's candidates: the instance pass at [thresholds[0], thresholds[bins - 1]] (camera cut). */
  or_view band = *v;
  band.error_threshold = thresholds[bins - 1];
  c.v = &band;
  load_bands(&c);
  float camera_scale = get_state(&c, S_SCALE);
  float high = thresholds[bins - 1] * camera_scale;
  float low = thresholds[0] * camera_scale;
  c.low[0] = low;
  c.high[0] = high;
  cells_pass(&c);
  float cache[MAX_LEVELS * 8];
  for (uint32_t k = 0; k < counters[0]; k++) {
    const uint32_t* cell = &c.cell_list[k * 4];
    uint32_t slot = cell[0];
    const float* levers = mesh_record(s, slot, 3);
    uint32_t level_count = (uint32_t)levers[2];
    if (level_count > MAX_LEVELS) level_count = MAX_LEVELS;
    memcpy(cache, &s->mesh_table[(uint32_t)levers[3] * 4], sizeof(float) * level_count * 8);
    const float* object_sphere = mesh_record(s, slot, 0);
    const float* lod_bounds = mesh_record(s, slot, 1);
    const float* flags = mesh_record(s, slot, 5);
    int occlusion_on = flags[2] > 0.5f;
    float min_pixels = fmaxf(v->min_pixel_radius, levers[1]);
    for (uint32_t i = cell[1]; i < cell[1] + cell[2]; i++) {
      const float* m = &s->instances[(size_t)i * INSTANCE_FLOATS];
      float scale = model_scale(m);
      f4 center = mat_mul(m, point(object_sphere));
      float radius = object_sphere[3] * scale;
      if (!(v->frustum_culling == 0 || frustum_distance(v, center.x, center.y, center.z, 0, 0) >= -radius)) continue;
      f4 vc = mat_mul(v->view, (f4){center.x, center.y, center.z, 1.0f});
      float view_distance = len3(vc.x, vc.y, vc.z);
      if (view_distance > radius && (radius * v->proj_scale) / view_distance < min_pixels) continue;
      if ((view_distance - radius) > levers[0]) continue;
      if (occluded_sphere(&c, occlusion_on, center.x, center.y, center.z, radius)) continue;
      f4 lod_center = mat_mul(v->view, mat_mul(m, point(lod_bounds)));
      float d = len3(lod_center.x, lod_center.y, lod_center.z);
      float r = lod_bounds[3] * scale;
      float pmin = (scale * v->proj_scale) / fmaxf(d - r, v->near_plane);
      float pmax = (scale * v->proj_scale) / fmaxf(d + r, v->near_plane);
      uint32_t first = 0xffffffffu, last = 0;
      for (uint32_t l = 0; l < level_count; l++) {
        const float* l0 = &cache[l * 8];
        if ((l0[2] * pmax <= high) && (l0[3] >= K_INF_HALF || l0[3] * pmin > low)) {
          uint32_t lf = (uint32_t)l0[0], size = (uint32_t)l0[1];
          if (lf < first) first = lf;
          if (lf + size > last) last = lf + size;
        }
      }
      for (uint32_t meshlet = first; meshlet < last; meshlet++) {
        const float* b = &s->bounds[(size_t)meshlet * 16];
        const float* cull = b + 12;
        f4 cc = mat_mul(m, point(cull));
        if (!(v->frustum_culling == 0 || frustum_distance(v, cc.x, cc.y, cc.z, 0, 0) >= -(cull[3] * scale))) continue;
        if (occluded_sphere(&c, occlusion_on, cc.x, cc.y, cc.z, cull[3] * scale)) continue;
        float own = projected_error(v, m, scale, b, b[8]);
        float parent = projected_error(v, m, scale, b + 4, b[9]);
        uint32_t a = lower_bound_scaled(thresholds, bins, camera_scale, own);
        uint32_t e = lower_bound_scaled(thresholds, bins, camera_scale, parent);
        if (a < e) {
          uint32_t tris = s->info[meshlet * 4];
          diff[a] += tris;
          diff[e] -= tris;
        }
      }
    }
  }
  uint64_t run = 0;
  for (uint32_t b = 0; b < bins; b++) {
    run += diff[b];
    triangles[b] = run;
  }
  free(diff);
  free(counters);
  free(c.cell_list);
  return 0;
}

/*
 * expandPass(cut), GeometryPool.ts:997-1021: for draw-list positions at < C_DRAWN of `cut`, the entry's
 * mesh slot is its instance record's .w of column 4, its part comes from the blend bits, and triangle i
 * writes 3 indices at regionStart + firstIndex + 3i, each (position << 7) | local id (GeometryPool.ts:92,
 * 1009), when index + 2 < capacity. Sums are u32 as in WGSL; an index whose + 2 wraps past 2^32 is
 * skipped (the reference's storage access would be out of bounds there).
 */
int or_expand(const or_scene* s, const uint32_t* triangles, uint64_t triangle_words, int cut, const uint32_t* counters,
              const uint32_t* draw_list, uint32_t first, uint32_t* indices, uint64_t* written) {
  if (!s || !triangles || !counters || !draw_list || !indices || cut < 0 || cut > 1) return -1;
  const uint32_t capacity = s->cut_capacity[cut];
  const uint32_t drawn = counters[2 + cut * CUT_COUNTERS + C_DRAWN];
  uint64_t n = 0;
  for (uint32_t at = first; at < drawn; at++) {
    const uint32_t position = at + (uint32_t)cut * s->max_meshlets;
    const uint32_t* e = &draw_list[(size_t)position * 4];
    if (e[0] >= s->instance_count || e[1] >= s->meshlet_count) return -2;
    const uint32_t slot = (uint32_t)s->instances[(size_t)e[0] * INSTANCE_FLOATS + 19];
    const uint32_t part = (e[3] & 0xffff00u) == FULL_FADE_BITS ? 0u : 1u;
    const uint32_t region = counters[PER_MESH_COUNTERS + (uint32_t)cut * s->slots * 4 + 1 + slot * 4 + part * 2];
    const uint32_t first_triangle = s->info[(size_t)e[1] * 4 + 1] / 3u;
    const uint32_t base = position << 7;
    const uint32_t count = e[3] & 255u;
    for (uint32_t i = 0; i < count; i++) {
      const uint32_t index = region + e[2] + i * 3u;
      if (index > 0xfffffffdu || index + 2u >= capacity) continue;
      if ((uint64_t)first_triangle + i >= triangle_words) return -3;
      const uint32_t packed = triangles[first_triangle + i];
      indices[index] = base | (packed & 255u);
      indices[index + 1] = base | ((packed >> 8) & 255u);
      indices[index + 2] = base | ((packed >> 16) & 255u);
      n += 3;
    }
  }
  if (written) *written = n;
  return 0;
}

/* ------------------------------------------------------------------ occluder raster (vg_hzb_from_cut) */

typedef struct {
  int32_t x, y; /* 1/256 pixel, y down */
  float z;
  int ok;
} rvert_t;

static rvert_t raster_vertex(const float* m, const float* vp, const float* p, uint32_t W, uint32_t H) {
  rvert_t v = {0, 0, 0.0f, 0};
  const float wx = ((m[0] * p[0] + m[4] * p[1]) + m[8] * p[2]) + m[12];
  const float wy = ((m[1] * p[0] + m[5] * p[1]) + m[9] * p[2]) + m[13];
  const float wz = ((m[2] * p[0] + m[6] * p[1]) + m[10] * p[2]) + m[14];
  const float cx = ((vp[0] * wx + vp[4] * wy) + vp[8] * wz) + vp[12];
  const float cy = ((vp[1] * wx + vp[5] * wy) + vp[9] * wz) + vp[13];
  const float cz = ((vp[2] * wx + vp[6] * wy) + vp[10] * wz) + vp[14];
  const float cw = ((vp[3] * wx + vp[7] * wy) + vp[11] * wz) + vp[15];
  if (!(cw > 0.0f) || !(cz >= 0.0f)) return v;
  const float sx = ((cx / cw) * 0.5f + 0.5f) * (float)W;
  const float sy = (0.5f - (cy / cw) * 0.5f) * (float)H;
  if (!(fabsf(sx) <= 1048576.0f) || !(fabsf(sy) <= 1048576.0f)) return v;
  v.x = (int32_t)rintf(sx * 256.0f);
  v.y = (int32_t)rintf(sy * 256.0f);
  v.z = cz / cw;
  v.ok = 1;
  return v;
}

static int32_t floor_shift8(int64_t v) { return (int32_t)(v >= 0 ? v / 256 : -((-v + 255) / 256)); }

static void raster_triangle(rvert_t a, rvert_t b, rvert_t c, uint32_t W, uint32_t H, float* depth) {
  if (!a.ok || !b.ok || !c.ok) return;
  const int64_t signed_area = (int64_t)(b.x - a.x) * (c.y - a.y) - (int64_t)(b.y - a.y) * (c.x - a.x);
  if (signed_area >= 0) return; /* back-facing (screen y is down) or degenerate */
  const rvert_t t = b;           /* edge order: positive area */
  b = c;
  c = t;
  const int64_t area = -signed_area;
  int32_t minx = a.x, maxx = a.x, miny = a.y, maxy = a.y;
  if (b.x < minx) minx = b.x;
  if (c.x < minx) minx = c.x;
  if (b.x > maxx) maxx = b.x;
  if (c.x > maxx) maxx = c.x;
  if (b.y < miny) miny = b.y;
  if (c.y < miny) miny = c.y;
  if (b.y > maxy) maxy = b.y;
  if (c.y > maxy) maxy = c.y;
  int32_t x0 = floor_shift8((int64_t)minx - 128 + 255), x1 = floor_shift8((int64_t)maxx - 128);
  int32_t y0 = floor_shift8((int64_t)miny - 128 + 255), y1 = floor_shift8((int64_t)maxy - 128);
  if (x0 < 0) x0 = 0;
  if (y0 < 0) y0 = 0;
  if (x1 > (int32_t)W - 1) x1 = (int32_t)W - 1;
  if (y1 > (int32_t)H - 1) y1 = (int32_t)H - 1;
  for (int32_t y = y0; y <= y1; y++)
    for (int32_t x = x0; x <= x1; x++) {
      const int64_t px = (int64_t)x * 256 + 128, py = (int64_t)y * 256 + 128;
      const int64_t e0 = (int64_t)(c.x - b.x) * (py - b.y) - (int64_t)(c.y - b.y) * (px - b.x);
      const int64_t e1 = (int64_t)(a.x - c.x) * (py - c.y) - (int64_t)(a.y - c.y) * (px - c.x);
      const int64_t e2 = (int64_t)(b.x - a.x) * (py - a.y) - (int64_t)(b.y - a.y) * (px - a.x);
      if (e0 < 0 || e1 < 0 || e2 < 0) continue;
      float z = (((float)e0 * a.z + (float)e1 * b.z) + (float)e2 * c.z) / (float)area;
      z = fminf(fmaxf(z, 0.0f), 1.0f);
      float* d = &depth[(size_t)y * W + (size_t)x];
      if (z < *d) *d = z;
    }
}

int or_raster(const float* instances, uint32_t instance_count, const uint32_t* info, uint32_t meshlet_count,
              const uint32_t* triangles, uint64_t triangle_words, const uint32_t* meshlet_vertices,
              uint64_t meshlet_vertex_words, const float* vertices, uint64_t vertex_count, const uint32_t* draw_list,
              uint32_t count, const float* view_proj, uint32_t width, uint32_t height, float* depth) {
  rvert_t v[128];
  for (uint32_t e = 0; e < count; e++) {
    const uint32_t instance = draw_list[e * 4], meshlet = draw_list[e * 4 + 1];
    if (instance >= instance_count || meshlet >= meshlet_count) return -1;
    const uint32_t* mi = &info[(size_t)meshlet * 4];
    if (mi[3] == 0xffffffffu) continue; /* the mesh was added without positions */
    const uint64_t first = mi[1] / 3;
    if (first + mi[0] > triangle_words) return -1;
    uint32_t nv = 0;
    for (uint32_t t = 0; t < mi[0]; t++)
      for (int k = 0; k < 3; k++) {
        const uint32_t l = (triangles[first + t] >> (8 * k)) & 255u;
        if (l + 1 > nv) nv = l + 1;
      }
    if (nv > 128 || (uint64_t)mi[3] + nv > meshlet_vertex_words) return -1;
    const float* m = &instances[(size_t)instance * INSTANCE_FLOATS];
    for (uint32_t k = 0; k < nv; k++) {
      const uint32_t id = meshlet_vertices[mi[3] + k];
      if (id >= vertex_count) return -1;
      v[k] = raster_vertex(m, view_proj, &vertices[(size_t)id * 4], width, height);
    }
    for (uint32_t t = 0; t < mi[0]; t++) {
      const uint32_t packed = triangles[first + t];
      raster_triangle(v[packed & 255u], v[(packed >> 8) & 255u], v[(packed >> 16) & 255u], width, height, depth);
    }
  }
  return 0;
}

/* ---------------------------------------------------------------- test hooks */

/* The capacity controller on given counts (no scene): checkNode (steer 0) or steerNode (steer 1) for both
   cuts, GeometryPool.ts:1027-1069, updating the 8 state words; redo[cut] = whether checkNode redoes it. */
int or_test_controller(int steer, const uint32_t* requested, const uint32_t* selected, const uint32_t* capacity,
                       uint32_t max_meshlets, float lod_scale_limit, uint32_t* state, int* redo) {
  or_scene sc;
  or_view v;
  or_stats st;
  ctx_t c;
  memset(&sc, 0, sizeof(sc));
  memset(&v, 0, sizeof(v));
  memset(&st, 0, sizeof(st));
  memset(&c, 0, sizeof(c));
  uint32_t counters[PER_MESH_COUNTERS + 8];
  memset(counters, 0, sizeof(counters));
  sc.slots = 1;
  sc.mesh_slots = 1;
  sc.max_meshlets = max_meshlets;
  sc.cut_capacity[0] = capacity[0];
  sc.cut_capacity[1] = capacity[1];
  v.lod_scale_limit = lod_scale_limit;
  c.sc = &sc;
  c.v = &v;
  c.st = &st;
  c.state = state;
  c.counters = counters;
  for (int cut = 0; cut < 2; cut++) {
    *counter(&c, cut, C_REQUESTED) = requested[cut];
    *counter(&c, cut, C_SELECTED) = selected[cut];
  }
  if (steer) {
    steer_pass(&c);
  } else {
    check_pass(&c);
    for (int cut = 0; cut < 2; cut++) redo[cut] = (int)st.retries[cut];
  }
  return 0;
}

/* prefixPass (GeometryPool.ts:973-995) of one cut on given per-(slot, part) index counts (2 per slot,
   slot-major), from region `offset`: fills `counters` (or_counter_words(slots) words; the pairs' starts,
   C_REQUESTED, C_DRAWN from `selected`) and that cut's draw args in `args` (2 * slots * 2 * 5 words). */
int or_test_prefix(uint32_t slots, uint32_t mesh_slots, int cut, const uint32_t* counts, uint32_t capacity, uint32_t selected,
                   uint32_t max_meshlets, uint32_t offset, uint32_t* counters, uint32_t* args) {
  if (mesh_slots > slots || cut < 0 || cut > 1) return -1;
  or_scene sc;
  or_stats st;
  ctx_t c;
  memset(&sc, 0, sizeof(sc));
  memset(&st, 0, sizeof(st));
  memset(&c, 0, sizeof(c));
  sc.slots = slots;
  sc.mesh_slots = mesh_slots;
  sc.max_meshlets = max_meshlets;
  sc.cut_capacity[cut] = capacity;
  c.sc = &sc;
  c.st = &st;
  c.counters = counters;
  c.args = args;
  memset(counters, 0, sizeof(uint32_t) * or_counter_words(slots));
  for (uint32_t i = 0; i < mesh_slots; i++)
    for (uint32_t part = 0; part < 2; part++) *mesh_count(&c, cut, i, part) = counts[i * 2 + part];
  *counter(&c, cut, C_SELECTED) = selected;
  prefix_pass_from(&c, cut, offset);
  return 0;
}
