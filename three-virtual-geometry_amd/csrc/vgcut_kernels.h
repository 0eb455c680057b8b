// Device side of the cut: binary32 helpers in the reference kernels' expression order, the launch
// parameter block and the per-view records the kernels read. Compiled with -ffp-contract=off and
// correctly rounded division and square root, so every expression below rounds exactly as written.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vg {

// GeometryPool.ts:96-140, constants.ts:2-20, lodBlend.ts:22-24
constexpr uint32_t kMeshStride = 6;      // float4 per mesh record
constexpr uint32_t kMaxLevels = 48;      // level-table entries read per mesh
constexpr uint32_t kCellSize = 128;      // instances per cell
constexpr uint32_t kCutCounters = 5;
constexpr uint32_t kSelected = 0, kDrawn = 1, kSurvived = 2, kWork = 3, kRequested = 4;
constexpr uint32_t kPerMeshCounters = 12;
constexpr uint32_t kStateScale = 0, kStateFill = 2, kStateRetry = 4, kStateWords = 8;
constexpr uint32_t kFastPathMax = 4;
constexpr uint32_t kFadeSteps = 255;
constexpr uint32_t kFullFadeBits = 255u << 16;
constexpr uint32_t kHzbMaxLevels = 16;
constexpr uint32_t kHzbHeader = kHzbMaxLevels * 4 + 4;
constexpr uint32_t kHzbLevelCountSlot = kHzbMaxLevels * 4;
// Work-queue counter: item count in the high 28 bits, task (cluster test) count in the low 36 bits.
constexpr int kTaskBits = 36;
constexpr uint64_t kTaskMask = (1ull << kTaskBits) - 1;
constexpr uint64_t kTaskLimit = 1ull << 35;
constexpr uint32_t kMaxWorkItems = (1u << 28) - 1;
// The work list is kWorkParts partitions, each with its own packed counter (kWorkCounterSpacing u64
// apart, so on separate cache lines): waves of the instance pass append to the partition of their rank,
// which spreads the appends over kWorkParts atomics instead of one. The meshlet pass walks the
// partitions' task ranges one after another as one flat task space.
constexpr uint32_t kWorkParts = 16;
// Launches of at least this many views split the work list by view (partition = view % kWorkParts).
constexpr uint32_t kViewPartsMin = kWorkParts;
constexpr uint32_t kWorkCounterSpacing = 32;
// Metrics words (u64) per view slot.
constexpr uint32_t kMetricWords = 16;
enum : uint32_t {
  M_CELLS = 0, M_INSTANCES = 1, M_WORK0 = 2, M_WORK1 = 3, M_CLUSTERS0 = 4, M_CLUSTERS1 = 5, M_OCCLUSION = 6,
  M_BUDGET_TESTS = 7, M_THRESHOLD = 8, M_BUDGET_BIN = 9
};
constexpr uint32_t kMaxBudgetBins = 1024;
// k_budget_meshlets keeps one histogram per launched view in its block's LDS (32-bit partial sums, flushed by
// exchange) when they fit in this many words; otherwise each wave keeps its current view's histogram.
constexpr uint32_t kBudgetBlockHistWords = 10240;  // 40 KB: 39 views at the default 256 bins (with the static LDS, under 64 KB)
// A wave flushes its block's partial sums after adding this many meshlets (of at most 255 triangles each):
// between two flushes, each of the four waves adds at most 2^20 + 64 meshlets, so any one word moves by less
// than 4 * (2^20 + 64) * 255 < 2^31.
constexpr uint32_t kBudgetFlushMeshlets = 1u << 20;
// A LOD run: consecutive meshlets of one level whose LOD test has the same inputs (same LOD and parent
// spheres and errors; a zero error projects to 0 and an error >= 5e29 to 1e30 whatever the sphere). The
// meshlet pass tests each run once. Runs hold at most kMaxRun meshlets.
constexpr uint32_t kMaxRun = 4;
// Work-item tag: mesh slot in bits 0-15, then flags, then the launched view's index.
constexpr uint32_t kTagSlotMask = 0xffffu;
constexpr uint32_t kTagShadow = 1u << 16;     // the item belongs to the shadow cut
constexpr uint32_t kTagOcclusion = 1u << 17;  // the mesh opted into occlusion (camera cut only)
constexpr uint32_t kTagFast = 1u << 18;       // fast path: tasks are meshlets drawn without cluster tests
constexpr int kTagViewShift = 19;
constexpr uint32_t kMaxLaunchViews = 1u << (32 - kTagViewShift);
// LOD blocks: up to kBlockRuns consecutive LOD runs of one level. The meshlet pass first tests a block as
// a whole against bounds of its runs (a sphere around their LOD-sphere centres and one around their parent
// centres, the smallest own error and LOD radius, the largest parent error and parent radius) and skips it
// when every run's own error certainly projects above the band or every parent's certainly below it; the
// runs of the other blocks are tested one by one, exactly. Items of the meshlet pass are block ranges.
constexpr uint32_t kBlockRuns = 16;
// Per-wave LDS queue of blocks that passed, awaiting their runs' tests: at most 63 entries with untested
// runs plus the last consumed one, plus one chunk's 64 survivors.
constexpr uint32_t kBlockQueue = 128;
// Per-wave LDS ring of meshlets in the cut awaiting their visibility test: a wave appends its lanes' k-th
// meshlets one k at a time (at most 64 each) and drains full waves in between, so at most 63 + 64 wait.
constexpr uint32_t kQueue = 128;
static_assert((kQueue & (kQueue - 1)) == 0, "ring positions wrap by masking");
// An index is (draw-list position << kLocalVertexBits) | local vertex (GeometryPool.ts:92, 1009-1018).
constexpr uint32_t kLocalVertexBits = 7;  // ceil(log2(MAX_MESHLET_VERTICES = 128))
constexpr uint32_t kExpandBlocksPerCu = 8;  // persistent 256-thread blocks of the expand pass
// Per view slot words of two-phase occlusion: meshlet rejects recorded, whether more were dropped, and
// the camera band of the cut's final round (binary32 bits), which the second phase tests with.
constexpr uint32_t kPhaseWords = 8;
enum : uint32_t { P_REJECTS = 0, P_OVERFLOW = 1, P_LOW = 2, P_HIGH = 3, P_FADE = 4 };

// Literals as the reference's WGSL sees them: JS doubles rounded to binary32.
#define VG_INF_HALF ((float)(1e30 * 0.5))
#define VG_INF ((float)1e30)
#define VG_MARGIN_LO ((float)(1 - 1e-3))
#define VG_MARGIN_HI ((float)(1 + 1e-3))

// Per launched view, in device memory. The first part is copied from vg_view; `low`, `high`,
// `fade_scale` (loadBands, GeometryPool.ts:652-665) are derived on the device from the slot's state.
struct ViewDev {
  float view[16];
  float proj[16];
  float planes[24];
  float sweep[4];
  float proj_scale, near_plane, error_threshold, shadow_error_scale;
  float lod_blend_max, lod_scale_limit, min_pixel_radius, max_error_threshold;
  uint32_t frustum, occlusion, slot, budget;
  float range_lo, range_hi;
  uint32_t hzb_valid;
  uint32_t fresh;  // 1: the slot's controller state starts at rest this launch (scale 1, as on a new slot)
  uint32_t two_phase;  // 1: occlusion in two phases: the cut records its occlusion rejects for the second
  // derived
  float low[2], high[2];
  float fade_scale;
  uint32_t redo[2];  // retry flags (S_RETRY > 0.5) of the current round
  uint32_t budget_bin;
  float budget_scale, budget_low, budget_high;  // camera scale and band of the budget histogram
  float budget_inv;  // (bins - 1) / ln(T[B-1] / T[0]): estimates a value's bin in the log-spaced table
  float sweep_reach[6];  // per frustum plane: fmaxf(plane normal . sweep, 0), the swept frustum's extra reach
  // Bounds on the view matrix for the block test: its 3 x 3 part's spectral norm (Gershgorin on V^T V) and
  // largest row sum of magnitudes, and its largest translation magnitude.
  float norm_bound, row_bound, move_bound;
  float frustum_n2, frustum_c;  // bounds on the frustum planes (frustum_bounds)
};

struct Params {
  // scene
  const float4* __restrict__ mesh_table;  // slot records then level tables (GeometryPool.ts:432-460)
  const float4* __restrict__ m_cull;      // per meshlet: culling sphere
  const float4* __restrict__ m_lod;       // LOD sphere
  const float4* __restrict__ m_parent;    // parent LOD sphere
  const float2* __restrict__ m_err;       // (error, parentError)
  const uint32_t* __restrict__ m_tris;    // triangle count
  const uint32_t* __restrict__ m_run;     // LOD run of each meshlet
  const uint2* __restrict__ r_span;       // per run: (first meshlet, meshlet count)
  const float4* __restrict__ r_lod;       // per run: LOD sphere
  const float4* __restrict__ r_parent;    // parent LOD sphere
  const float2* __restrict__ r_err;       // (error, parentError)
  const uint32_t* __restrict__ r_block;   // block of each run
  const uint2* __restrict__ level_blocks; // per level (mesh-table level entry pair): block of its first meshlet, block of its last + 1
  const uint2* __restrict__ b_span;       // per block: (first run, run count)
  const float4* __restrict__ b_lod;       // per block: centre and radius of a sphere holding its runs' LOD-sphere centres
  const float4* __restrict__ b_par;       // the same for their parent spheres' centres
  const float4* __restrict__ b_err;       // (smallest own error, smallest LOD radius, largest parent error or +inf, largest parent radius)
  const float4* __restrict__ i_col0;      // instance matrix columns, SoA
  const float4* __restrict__ i_col1;
  const float4* __restrict__ i_col2;
  const float4* __restrict__ i_col3;
  const float4* __restrict__ c_sphere;    // per cell: bounding sphere
  const float4* __restrict__ c_rec;       // (maxRadius, slot, firstInstance, count)
  const float4* __restrict__ c_bound;     // (A, R_in, R_cull, 0): bounds of the cell's instances (cell_inside_bits)
  uint32_t cell_count;
  uint32_t mesh_slots;   // highest used slot + 1
  uint32_t slots;        // slot capacity
  uint32_t max_meshlets;
  uint32_t cut_capacity[2];
  // per view slot
  uint32_t* counters;    // counter_stride words per slot (state words at the end)
  uint32_t counter_stride;
  uint32_t* args;        // args_stride words per slot
  uint32_t args_stride;
  uint4* const* draws;   // per slot: 2 * max_meshlets uvec4
  float* const* hzbs;    // per slot HZB buffer
  unsigned long long* metrics;  // kMetricWords per slot
  // launch
  ViewDev* views;
  uint32_t view_count;
  uint2* cell_queue;     // (view, cell)
  uint32_t* cell_queue_count;
  uint4* work;           // (instance, first, last, tag)
  unsigned long long* work_start;  // first task of each item
  float* work_scale;     // instance scale of each item (max column length, computed once)
  unsigned long long* work_counter;  // kWorkParts packed (items << kTaskBits | tasks) counters
  uint32_t work_capacity;            // items per partition; partition q holds items [q * work_capacity, ...)
  uint32_t* flags;       // [0], [1] a view needs retry round 1, 2; [2] internal error; [3] work-list overflow;
                         // [4] most work items any pass appended (sizes later work lists)
  // footprint (measurement only): bit per instance tested / per meshlet in a candidate range
  uint32_t* fp_instances;
  uint32_t* fp_meshlets;
  // triangle budget
  const float* budget_thresholds;  // bins, ascending
  unsigned long long* budget_hist; // per view: bins + 1
  uint32_t budget_bins;
  uint32_t budget_block_hist;      // 1: k_budget_meshlets keeps every launched view's histogram in its block's LDS
  // expand pass (GeometryPool.ts:997-1021)
  const uint32_t* __restrict__ i_slot;   // mesh slot of each instance
  const uint32_t* __restrict__ m_first;  // meshletInfo.y: first index of each meshlet's triangles (pool)
  const uint32_t* __restrict__ m_tri;    // meshletTriangles: local ids a | b << 8 | c << 16
  uint32_t* const* indices;              // per slot: camera index buffer, shadow index buffer
  unsigned long long* expand_plan;       // per (view, cut): first 64-entry chunk; then the chunk total
  uint4* expand_records;                 // per drawn entry (64 per chunk): first index, first triangle, triangles, position
  // Two-phase occlusion: a cut records, for views that ask for it, the camera cut's occlusion rejects of
  // its final round; the second phase re-tests them against the slot's new HZB.
  uint32_t phase;                        // 0: a cut (recording for two-phase views); 2: the second phase
  uint32_t* reject_bits;                 // per view slot: reject_words words; bit i: instance i occluded
  uint32_t reject_words;
  uint4* reject_meshlets;                // per view slot: (instance, meshlet, fades | mesh slot << 16, scale bits)
  uint32_t reject_capacity;              // meshlet rejects per view slot
  uint32_t* phase_words;                 // per view slot: kPhaseWords (see P_*)
  const uint32_t* counters1;             // second phase: the cut's counters (p.counters holds the second phase's)
};

__host__ __device__ inline uint32_t state_base(uint32_t slots) { return kPerMeshCounters + 2u * slots * 4u; }

// ------------------------------------------------------------------ arithmetic

__host__ __device__ inline float4 mat_mul(const float4& c0, const float4& c1, const float4& c2, const float4& c3, float x,
                                          float y, float z, float w) {
  float4 r;
  r.x = ((c0.x * x + c1.x * y) + c2.x * z) + c3.x * w;
  r.y = ((c0.y * x + c1.y * y) + c2.y * z) + c3.y * w;
  r.z = ((c0.z * x + c1.z * y) + c2.z * z) + c3.z * w;
  r.w = ((c0.w * x + c1.w * y) + c2.w * z) + c3.w * w;
  return r;
}

// base[index] through a 32-bit byte offset from a wave-uniform base (the SGPR-base form of a global load, no
// per-lane 64-bit address arithmetic). For arrays under 4 GB: a scene holds at most 2^24 meshlets, runs and
// instances, and a work-list partition at most 2^24 items.
template <typename T>
__device__ __forceinline__ T ld32(const T* __restrict__ base, uint32_t index) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + index * static_cast<uint32_t>(sizeof(T)));
}

template <typename P>
__host__ __device__ inline float4 mat_mul16(P m, float x, float y, float z, float w) {
  float4 r;
  r.x = ((m[0] * x + m[4] * y) + m[8] * z) + m[12] * w;
  r.y = ((m[1] * x + m[5] * y) + m[9] * z) + m[13] * w;
  r.z = ((m[2] * x + m[6] * y) + m[10] * z) + m[14] * w;
  r.w = ((m[3] * x + m[7] * y) + m[11] * z) + m[15] * w;
  return r;
}

// sqrt and / are correctly rounded: HIP's default -fhip-fp32-correctly-rounded-divide-sqrt lowers
// llvm.sqrt.f32 and fdiv to IEEE-exact sequences on gfx950 (checked against the host by a GPU test).
__host__ __device__ inline float len3(float x, float y, float z) { return __builtin_sqrtf((x * x + y * y) + z * z); }
template <typename P>
__host__ __device__ inline float dot3(P n, float x, float y, float z) { return (n[0] * x + n[1] * y) + n[2] * z; }
__host__ __device__ inline float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// frustumDistance, GeometryPool.ts:690-698. mode 0: plain, 1: swept (x 1), 2: swept x select(flag, 1, 0).
template <typename V>
__host__ __device__ inline float frustum_distance(const V& v, float cx, float cy, float cz, int mode, bool flag) {
  float best = 1e30f;
#pragma unroll
  for (int p = 0; p < 6; p++) {
    const auto n = &v.planes[p * 4];
    float d = dot3(n, cx, cy, cz) + n[3];
    // sweep_reach[p] = fmaxf(n . sweep, 0), computed once per view by k_begin (x 1.0f is exact).
    if (mode == 1) d = d + v.sweep_reach[p];
    if (mode == 2) d = d + v.sweep_reach[p] * (flag ? 1.0f : 0.0f);
    best = fminf(best, d);
  }
  return best;
}

// Bounds on a view's frustum planes for the inside test below: n2 = the largest |n|_2 of a plane normal
// (at least 1; its binary32 square root, a few ulp, covered by 1 + 1e-6), c = the largest |constant| plus n2
// times |sweep|_1 (so at least any plane's |constant| + |sweep_reach|). Computed once per view by k_begin.
template <typename V>
__host__ __device__ inline void frustum_bounds(V& v) {
  float n2 = 1.0f, c = 0.0f;
  for (int p = 0; p < 6; p++) {
    const float* n = &v.planes[p * 4];
    n2 = fmaxf(n2, __builtin_sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]) * 1.000001f);
    c = fmaxf(c, fabsf(n[3]));
  }
  v.frustum_n2 = n2;
  v.frustum_c = c + n2 * ((fabsf(v.sweep[0]) + fabsf(v.sweep[1])) + fabsf(v.sweep[2]));
}

// Which whole-cell answers of the instance pass's frustum tests a cell settles for one view (queue-entry
// bits kQ*): kQInCamera when every instance's bounding sphere certainly passes the camera test
// (frustum_distance(centre, 0) >= -radius), kQInShadow the same for the swept test (mode 1), kQCullCamera /
// kQCullShadow when every instance's culling sphere is certainly inside (frustum_distance(centre, mode) >=
// cull radius, the fast path's condition). `bound` (per cell, built on the host in binary64 from the binary32
// matrices): x = A, the largest |c0_k x| + |c1_k y| + |c2_k z| + |c3_k| over instances, components and both
// sphere centres (what the kernel's mat_mul rounds against); y = R_in >= |c_i - C| and z = R_cull >=
// |cc_i - C| + cr_i around the cell centre C (binary32, as stored), for the exact (real) centres c_i, cc_i of the
// instances' bounding and culling spheres and the culling spheres' radii cr_i.
// With n2 >= |n|_2 (n2 >= 1) and D = the exact plane distance of C: the exact plane distance of c_i is at least
// D - |n| |c_i - C| >= D - n2 R_in, and the kernel's binary32 value differs from the exact one by at most
// sqrt(3) gamma_4 n2 A (rounding of the centre) + gamma_5 (|n|_1 |c_i|_inf + c) (the dot product), as D's
// computed value here by gamma_5 (|n|_1 |C|_inf + c), with |n|_1 <= sqrt(3) n2 and gamma_5 < 3.1e-7: in all
// under 1.1e-6 n2 (|C|_inf + R + A) + 6.2e-7 c. So a computed D >= n2 (1 + 1e-5) R_in + err, err = 1e-5
// (2 n2 (|C|_inf + R + A) + c), leaves every instance's computed value >= 0 >= -r_i (radii are >= 0). For
// the culling spheres, D >= n2 (1 + 1e-5) R_cull + err leaves each computed value above (1 + 1e-5) cr_i,
// which bounds the kernel's binary32 cull radius (a product and a square root: a few ulp). NaN anywhere
// gives no bits.
constexpr uint32_t kQueueViewMask = (1u << 28) - 1u;
constexpr uint32_t kQInCamera = 1u << 28, kQInShadow = 1u << 29, kQCullCamera = 1u << 30, kQCullShadow = 1u << 31;

template <typename V>
__host__ __device__ inline uint32_t cell_inside_bits(const V& v, const float4& sphere, const float4& bound, bool sweep) {
  const float mag = fmaxf(fabsf(sphere.x), fmaxf(fabsf(sphere.y), fabsf(sphere.z))) + fmaxf(bound.y, bound.z);
  const float err = 1e-5f * (2.0f * v.frustum_n2 * (mag + bound.x) + v.frustum_c);
  const float reach = v.frustum_n2 * 1.00001f;
  const float in_r = reach * bound.y + err, cull_r = reach * bound.z + err;
  uint32_t bits = 0;
  const float d0 = frustum_distance(v, sphere.x, sphere.y, sphere.z, 0, false);
  if (d0 >= in_r) bits |= kQInCamera;
  if (d0 >= cull_r) bits |= kQCullCamera;
  if (sweep) {
    const float d1 = frustum_distance(v, sphere.x, sphere.y, sphere.z, 1, false);
    if (d1 >= in_r) bits |= kQInShadow;
    if (d1 >= cull_r) bits |= kQCullShadow;
  }
  return bits;
}

// fdlibm e_logf.c (Sun Microsystems, 1993): x = 2^k (1 + f), log(1 + f) = f - s (f - R(z)),
// s = f / (2 + f). Basic operations only, so the host and device give the same bits.
__host__ __device__ inline float vg_logf(float x) {
  const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f, two25 = 3.355443200e+07f;
  const float Lg1 = 6.6666668653e-01f, Lg2 = 4.0000000596e-01f, Lg3 = 2.8571429849e-01f, Lg4 = 2.2222198546e-01f,
              Lg5 = 1.8183572590e-01f, Lg6 = 1.5313838422e-01f, Lg7 = 1.4798198640e-01f;
  uint32_t ix = __builtin_bit_cast(uint32_t, x);
  int32_t k = 0;
  if (ix >= 0x7f800000u) return x + x;
  if (ix < 0x00800000u) {
    if (ix == 0) return -__builtin_inff();
    k -= 25;
    x *= two25;
    ix = __builtin_bit_cast(uint32_t, x);
  }
  k += static_cast<int32_t>(ix >> 23) - 127;
  ix &= 0x007fffffu;
  uint32_t i = (ix + (0x95f64u << 3)) & 0x800000u;
  float m = __builtin_bit_cast(float, ix | (i ^ 0x3f800000u));
  k += static_cast<int32_t>(i >> 23);
  float f = m - 1.0f;
  float dk = static_cast<float>(k);
  float s = f / (2.0f + f);
  float z = s * s;
  float w = z * z;
  float t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  float t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  float R = t2 + t1;
  float hfsq = 0.5f * f * f;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

// projectedError, GeometryPool.ts:678-683. kAffine: the instance matrix's last row is (0, 0, 0, 1) and the
// sphere finite, so the transformed w is exactly ((0 x + 0 y) + 0 z) + 1 = 1 and is not computed.
template <bool kAffine = false, typename V>
__host__ __device__ inline float projected_error(const V& v, const float4& c0, const float4& c1, const float4& c2,
                                                 const float4& c3, float scale, const float4& s, float error) {
  float4 q = mat_mul(c0, c1, c2, c3, s.x, s.y, s.z, 1.0f);
  float4 vc = mat_mul16(v.view, q.x, q.y, q.z, kAffine ? 1.0f : q.w);
  float distance = fmaxf(len3(vc.x, vc.y, vc.z) - s.w * scale, v.near_plane);
  float pixels = ((error * scale) / distance) * v.proj_scale;
  return error >= VG_INF_HALF ? VG_INF : pixels;
}

// Bounds of the block test (see kBlockRuns): whether every run of a block is certainly out of the band.
// For a run i of the block with LOD centre p_i, |p_i - c| <= rho around the block centre c, so in view space
// |V M p_i - V M c| <= |V|_2 |M|_2 rho (both norms bounded above by Gershgorin), and the exact binary32
// distance d_i = max(|V (M p_i)| - r_i s, near) of the run's own test is at most
//   dmax = max(|V M c| + |V|_2 |M|_2 rho - r_min s + slack, near),
// where the slack, 1e-5 times the magnitude of every term the two evaluations add (the view-space
// coordinates' sizes W, the distance, the reach, the radius), covers the rounding of both the exact
// path's products (at most 2 sqrt(3) gamma_4 W, about 8.3e-7 W, with no FMA) and this bound's own. Then
// own_i = e_i s P / d_i >= e_min s P / dmax, and e_min s P > high dmax (1 + 1e-5) leaves every own_i above
// high. Likewise with the parent centres, a lower bound dpmin on every parent distance, and
// pe_max s P (1 + 1e-5) <= low dpmin leaves every parent_i at most low (pe_max is +inf when a parent error
// is >= 5e29, which projects to 1e30). Only for affine scenes (the transformed w is exactly 1).
struct BlockBounds {
  float scale, norm_m, row_m, move_m;  // item scale, |M|_2 bound, largest row sum of |M|, largest |translation|
};

__host__ __device__ inline BlockBounds instance_bounds(const float4& c0, const float4& c1, const float4& c2, const float4& c3, float scale) {
  const float g00 = c0.x * c0.x + c0.y * c0.y + c0.z * c0.z, g11 = c1.x * c1.x + c1.y * c1.y + c1.z * c1.z;
  const float g22 = c2.x * c2.x + c2.y * c2.y + c2.z * c2.z;
  const float g01 = fabsf(c0.x * c1.x + c0.y * c1.y + c0.z * c1.z), g02 = fabsf(c0.x * c2.x + c0.y * c2.y + c0.z * c2.z);
  const float g12 = fabsf(c1.x * c2.x + c1.y * c2.y + c1.z * c2.z);
  BlockBounds b;
  b.scale = scale;
  b.norm_m = __builtin_sqrtf(fmaxf(g00 + g01 + g02, fmaxf(g01 + g11 + g12, g02 + g12 + g22)));
  b.row_m = fmaxf(fabsf(c0.x) + fabsf(c1.x) + fabsf(c2.x), fmaxf(fabsf(c0.y) + fabsf(c1.y) + fabsf(c2.y), fabsf(c0.z) + fabsf(c1.z) + fabsf(c2.z)));
  b.move_m = fmaxf(fabsf(c3.x), fmaxf(fabsf(c3.y), fabsf(c3.z)));
  return b;
}

// View-space distance from the camera to M c (FMA arithmetic: a bound, covered by the slack) and the
// magnitude W of the view-space coordinates of any point within rho of c.
template <typename V>
__host__ __device__ inline float2 block_distance(const V& v, const float4& c0, const float4& c1, const float4& c2, const float4& c3,
                                                 const BlockBounds& ib, const float4& sphere) {
  const float qx = fmaf(c0.x, sphere.x, fmaf(c1.x, sphere.y, fmaf(c2.x, sphere.z, c3.x)));
  const float qy = fmaf(c0.y, sphere.x, fmaf(c1.y, sphere.y, fmaf(c2.y, sphere.z, c3.y)));
  const float qz = fmaf(c0.z, sphere.x, fmaf(c1.z, sphere.y, fmaf(c2.z, sphere.z, c3.z)));
  const auto m = v.view;  // the record may be in the constant address space (scalar loads)
  const float x = fmaf(m[0], qx, fmaf(m[4], qy, fmaf(m[8], qz, m[12])));
  const float y = fmaf(m[1], qx, fmaf(m[5], qy, fmaf(m[9], qz, m[13])));
  const float z = fmaf(m[2], qx, fmaf(m[6], qy, fmaf(m[10], qz, m[14])));
  const float extent = fmaxf(fabsf(sphere.x), fmaxf(fabsf(sphere.y), fabsf(sphere.z))) + sphere.w;
  const float w = fmaf(v.row_bound, fmaf(ib.row_m, extent, ib.move_m), v.move_bound);
  return make_float2(__builtin_sqrtf(fmaf(x, x, fmaf(y, y, z * z))), w);
}

template <typename V>
__host__ __device__ inline bool block_out_of_band(const V& v, const float4& c0, const float4& c1, const float4& c2, const float4& c3,
                                                  const BlockBounds& ib, const float4& lod, const float4& par, const float4& err,
                                                  float low, float high) {
  const float kSlack = 1e-5f, kMargin = 1.00001f;
  const float s = ib.scale, stretch = v.norm_bound * ib.norm_m;
  const float2 own = block_distance(v, c0, c1, c2, c3, ib, lod);
  const float reach = stretch * lod.w, rmin = err.y * s;
  const float dmax = fmaxf(own.x + reach - rmin + kSlack * (own.y + own.x + reach + rmin), v.near_plane);
  if (err.x * s * v.proj_scale > high * dmax * kMargin) return true;
  const float2 pd = block_distance(v, c0, c1, c2, c3, ib, par);
  const float preach = stretch * par.w, rpmax = err.w * s;
  const float dpmin = fmaxf(pd.x - preach - rpmax - kSlack * (pd.y + pd.x + preach + rpmax), v.near_plane);
  return err.z * s * v.proj_scale * kMargin <= low * dpmin;
}

// Upper bounds on the view matrix's stretch for the block test: sqrt of the largest Gershgorin row sum of
// V^T V (>= the spectral norm of V's 3 x 3 part), its largest row sum of magnitudes, and the largest
// translation magnitude. The products are rounded up by the block test's slack.
__host__ __device__ inline void view_bounds(ViewDev& v) {
  const float* m = v.view;
  float g[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) g[i][j] = fabsf(m[i * 4] * m[j * 4] + m[i * 4 + 1] * m[j * 4 + 1] + m[i * 4 + 2] * m[j * 4 + 2]);
  float gersh = 0.0f, rows = 0.0f;
  for (int i = 0; i < 3; i++) {
    gersh = fmaxf(gersh, g[i][0] + g[i][1] + g[i][2]);
    rows = fmaxf(rows, fabsf(m[i]) + fabsf(m[4 + i]) + fabsf(m[8 + i]));
  }
  v.norm_bound = sqrtf(gersh);
  v.row_bound = rows;
  v.move_bound = fmaxf(fabsf(m[12]), fmaxf(fabsf(m[13]), fabsf(m[14])));
}

// ------------------------------------------------------------------ scalar (wave-uniform) reads

// Read-only data at wave-uniform addresses goes through the scalar cache: a pointer in the constant
// address space makes such loads s_load into SGPRs instead of per-lane vector loads. The data may be
// written by an earlier kernel of the launch sequence (each dispatch starts with that cache invalidated),
// never by the kernel that reads it this way.
#define VG_CONST __attribute__((address_space(4)))
#define VG_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const VG_CONST T* scalar(const T* p) {
  return (const VG_CONST T*)(p);
}
// Vector-type records (HIP's vector types cannot be copied out of the constant address space whole).
__device__ __forceinline__ float4 scalar_f4(const float4* p, size_t i) {
  const VG_CONST float* f = (const VG_CONST float*)(p + i);
  return make_float4(f[0], f[1], f[2], f[3]);
}
__device__ __forceinline__ uint2 scalar_u2(const uint2* p, size_t i) {
  const VG_CONST uint32_t* u = (const VG_CONST uint32_t*)(p + i);
  return make_uint2(u[0], u[1]);
}

// ------------------------------------------------------------------ wave64 helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Rank of the calling wave with consecutive ranks on one XCD: workgroups are dispatched to the 8 XCDs
// round-robin, so block b runs on XCD b % 8. Giving each XCD a contiguous eighth of the work keeps
// the data a wave touches (one mesh's clusters, neighbouring instances) in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_wave_rank() {
  uint32_t b = blockIdx.x;
  if ((gridDim.x & 7u) == 0) b = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  return b * (blockDim.x / 64) + threadIdx.x / 64;
}
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

template <typename T>
__device__ __forceinline__ T uniform(T v) {
  return static_cast<T>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

// Inclusive prefix sum over the 64 lanes with DPP moves: Hillis-Steele within each row of 16 lanes
// (row_shr 1, 2, 4, 8; lanes shifted in from outside the row read 0), then row_bcast:15 adds row 0's
// total to row 1 and row 2's to row 3, and row_bcast:31 adds the total of rows 0-1 to rows 2-3.
template <int kCtrl, int kRowMask, bool kBoundZero>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, kRowMask, 0xf, kBoundZero));
}
__device__ __forceinline__ uint32_t wave_scan(uint32_t v) {
  v += dpp<0x111, 0xf, true>(v);   // row_shr:1
  v += dpp<0x112, 0xf, true>(v);   // row_shr:2
  v += dpp<0x114, 0xf, true>(v);   // row_shr:4
  v += dpp<0x118, 0xf, true>(v);   // row_shr:8
  v += dpp<0x142, 0xa, false>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp<0x143, 0xc, false>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}
__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t lane) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), static_cast<int>(lane)));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v >> 32)), static_cast<int>(lane)));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// The OR over the 64 lanes (wave-uniform): the same DPP scan with OR (shifted-in lanes read 0).
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  v |= dpp<0x111, 0xf, true>(v);
  v |= dpp<0x112, 0xf, true>(v);
  v |= dpp<0x114, 0xf, true>(v);
  v |= dpp<0x118, 0xf, true>(v);
  v |= dpp<0x142, 0xa, false>(v);
  v |= dpp<0x143, 0xc, false>(v);
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// The largest value over the 64 lanes (wave-uniform): the same DPP pattern with max (shifted-in lanes read 0).
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  v = max(v, dpp<0x111, 0xf, true>(v));
  v = max(v, dpp<0x112, 0xf, true>(v));
  v = max(v, dpp<0x114, 0xf, true>(v));
  v = max(v, dpp<0x118, 0xf, true>(v));
  v = max(v, dpp<0x142, 0xa, false>(v));
  v = max(v, dpp<0x143, 0xc, false>(v));
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// The sum over the 64 lanes (wave-uniform).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(wave_scan(v)), 63));
}

__device__ __forceinline__ uint64_t wave_scan64(uint64_t v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t t = __shfl_up(v, d, 64);
    if (static_cast<int>(lane_id()) >= d) v += t;
  }
  return v;
}

// Groups the active lanes of a wave by address (`addr` per lane) for an aggregated add: each lane learns
// its group's leader (the group's lowest lane) and its exclusive prefix in the group; each leader learns
// the group's total. kOne: every active lane adds 1.
template <bool kOne>
__device__ __forceinline__ void wave_group(const uint32_t* addr, uint32_t value, bool active, uint32_t& leader_of, uint32_t& prefix_of,
                                           uint32_t& leader_total) {
  uint64_t pending = __ballot(active);
  const uint32_t lane = lane_id();
  const uint32_t a_lo = static_cast<uint32_t>(reinterpret_cast<uint64_t>(addr));
  const uint32_t a_hi = static_cast<uint32_t>(reinterpret_cast<uint64_t>(addr) >> 32);
  leader_of = prefix_of = leader_total = 0;
  while (pending) {
    const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(a_lo), leader));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(a_hi), leader));
    const bool mine = ((pending >> lane) & 1ull) && a_lo == lo && a_hi == hi;
    const uint64_t group = __ballot(mine);
    uint32_t prefix, total;
    if (kOne) {  // the exclusive prefix counts the group's lanes below this one
      prefix = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(group >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(group), 0u));
      total = static_cast<uint32_t>(__popcll(group));
    } else {
      const uint32_t v = mine ? value : 0u;
      const uint32_t incl = wave_scan(v);
      prefix = incl - v;
      total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
    }
    if (lane == static_cast<uint32_t>(leader)) leader_total = total;
    if (mine) {
      leader_of = static_cast<uint32_t>(leader);
      prefix_of = prefix;
    }
    pending &= ~group;
  }
}

// atomicAdd of `value` for every lane with `active`, aggregated per distinct address: one atomic per
// address per wave, all issued by their leaders in one instruction after the grouping (a loop of
// returning atomics into one register would wait for each before the next). Returns the lane's old
// value (the group base plus its exclusive prefix). All 64 lanes must call it together.
template <bool kOne = false>
__device__ __forceinline__ uint32_t wave_atomic_add(uint32_t* addr, uint32_t value, bool active) {
  uint32_t leader, prefix, total;
  wave_group<kOne>(addr, value, active, leader, prefix, total);
  uint32_t base = 0;
  if (total) base = atomicAdd(addr, total);  // a leader's own address is its group's
  const uint32_t b = __shfl(base, static_cast<int>(leader), 64);
  return active ? b + prefix : 0u;
}

// One atomicAdd per wave on a wave-uniform address; returns each active lane's old value.
__device__ __forceinline__ uint32_t wave_atomic_add_uniform(uint32_t* addr, uint32_t value, bool active) {
  uint32_t v = active ? value : 0u;
  uint32_t incl = wave_scan(v);
  uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
  uint32_t base = 0;
  if (lane_id() == 0 && total) base = atomicAdd(addr, total);
  base = uniform(base);
  return base + incl - v;
}

// Keyed 64-bit counter add: one atomic per distinct address per wave.
__device__ __forceinline__ void wave_count_add_keyed(unsigned long long* addr, uint64_t value, bool active) {
  uint64_t pending = __ballot(active);
  const uint32_t lane = lane_id();
  const uint64_t a = reinterpret_cast<uint64_t>(addr);
  while (pending) {
    const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(a)), leader));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(a >> 32)), leader));
    uint64_t la = (static_cast<uint64_t>(hi) << 32) | lo;
    bool mine = ((pending >> lane) & 1ull) && a == la;
    uint64_t group = __ballot(mine);
    uint64_t incl = wave_scan64(mine ? value : 0ull);
    uint64_t total = __shfl(incl, 63, 64);
    if (lane == static_cast<uint32_t>(leader) && total) atomicAdd(reinterpret_cast<unsigned long long*>(la), total);
    pending &= ~group;
  }
}

// 64-bit counter add of a 32-bit per-lane value: one atomic per wave.
__device__ __forceinline__ void wave_count_add(unsigned long long* addr, uint32_t value, bool active) {
  const uint32_t total = wave_sum(active ? value : 0u);
  if (lane_id() == 0 && total) atomicAdd(addr, static_cast<unsigned long long>(total));
}

}  // namespace vg
