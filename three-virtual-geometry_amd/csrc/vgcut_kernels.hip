// Kernels of the per-frame cut (GeometryPool.ts:568-1092, OcclusionCulling.ts:119-170, 275-352).
//
// One launch sequence serves any number of views (the plain path is one view):
//   k_begin      zero the views' counters, derive the threshold bands from the controller state
//   k_cells      one thread per (cell, 16 views): cell culls, wave-compacted append to a (view, cell) queue
//   k_instances  one wave per queued (view, cell), 2 instances per lane: instance culls, occlusion,
//                level ranges, fast path; work items appended with one 64-bit atomic per wave that
//                returns both the item slot and the item's first cluster test, so the work list is
//                a prefix sum of cluster tests in item order
//   k_meshlets   persistent waves over equal shares of the flat cluster-test space: each wave finds
//                its first item by a 64-ary search and walks the items with wave shuffles, so short
//                and long ranges load-balance perfectly; drawn clusters are compacted with ballot and
//                one atomic per (view, cut) per wave, firstIndex with one atomic per (view, cut, slot, part)
//   k_prefix     one wave per (view, cut): region layout + indirect args; then the capacity check
//                (retry rounds) or the steer, one thread per view
// The hierarchical-Z pyramid is built by k_hzb_pass: 32 x 32 texel tiles reduced 5 levels at a time
// in LDS (max of 2 x 2, edge clamped).
#include "vgcut_kernels.h"

namespace vg {

__device__ __forceinline__ uint32_t* view_counters(const Params& p, uint32_t slot) {
  return p.counters + static_cast<size_t>(slot) * p.counter_stride;
}
__device__ __forceinline__ uint32_t* counter_at(const Params& p, uint32_t slot, uint32_t cut, uint32_t k) {
  return view_counters(p, slot) + 2 + cut * kCutCounters + k;
}
__device__ __forceinline__ uint32_t* mesh_count_at(const Params& p, uint32_t slot, uint32_t cut, uint32_t mesh, uint32_t part) {
  return view_counters(p, slot) + kPerMeshCounters + cut * p.slots * 4 + mesh * 4 + part * 2;
}
__device__ __forceinline__ uint32_t* state_at(const Params& p, uint32_t slot) {
  return view_counters(p, slot) + state_base(p.slots);
}
__device__ __forceinline__ unsigned long long* metric_at(const Params& p, uint32_t slot, uint32_t k) {
  return p.metrics + static_cast<size_t>(slot) * kMetricWords + k;
}

__device__ __forceinline__ uint32_t* phase_at(const Params& p, uint32_t slot) { return p.phase_words + static_cast<size_t>(slot) * kPhaseWords; }

// Whether any of instances [first, first + count) is a recorded occlusion reject of the slot.
__device__ __forceinline__ bool any_rejected(const Params& p, uint32_t slot, uint32_t first, uint32_t count) {
  if (count == 0) return false;
  const uint32_t* bits = p.reject_bits + static_cast<size_t>(slot) * p.reject_words;
  const uint32_t last = first + count - 1;
  for (uint32_t w = first >> 5; w <= last >> 5; w++) {
    const uint32_t lo = w == (first >> 5) ? (first & 31u) : 0u;
    const uint32_t hi = w == (last >> 5) ? (last & 31u) : 31u;
    const uint32_t mask = (hi == 31u ? 0xffffffffu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
    if (bits[w] & mask) return true;
  }
  return false;
}

// Records the occlusion rejects of instances [first, first + n) (n <= 64; bit j of `occluded`: instance
// first + j) in the slot's bitmap, replacing what an earlier round recorded for them. Lane 0 only.
__device__ __forceinline__ void record_instance_rejects(const Params& p, uint32_t slot, uint32_t first, uint32_t n, uint64_t occluded) {
  uint32_t* bits = p.reject_bits + static_cast<size_t>(slot) * p.reject_words;
  const uint32_t last = first + n - 1;
  for (uint32_t w = first >> 5; w <= last >> 5; w++) {
    const uint32_t lo = w == (first >> 5) ? (first & 31u) : 0u;
    const uint32_t hi = w == (last >> 5) ? (last & 31u) : 31u;
    const uint32_t range = (hi == 31u ? 0xffffffffu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
    // bits of `occluded` for this word: instance w * 32 + b is bit (w * 32 + b - first) of the mask
    const int shift = static_cast<int>(w * 32u) - static_cast<int>(first);
    const uint64_t m = shift >= 0 ? (occluded >> shift) : (occluded << (-shift));
    const uint32_t set = static_cast<uint32_t>(m) & range;
    atomicAnd(&bits[w], ~range);
    if (set) atomicOr(&bits[w], set);
  }
}

__device__ __forceinline__ float get_state(const uint32_t* state, uint32_t i) {
  return static_cast<float>(state[i]) / 10000.0f;
}
__device__ __forceinline__ void set_state(uint32_t* state, uint32_t i, float value) {
  state[i] = static_cast<uint32_t>(rintf(value * 10000.0f));
}

// loadBands, GeometryPool.ts:652-665 (WGSL smoothstep).
__device__ void load_bands(ViewDev& v, const uint32_t* state) {
  float camera_scale = get_state(state, kStateScale);
  float shadow_scale = get_state(state, kStateScale + 1);
  float x = v.error_threshold * camera_scale;
  float t = clampf((x - 1.25f) / (2.5f - 1.25f), 0.0f, 1.0f);
  float smooth = (t * t) * (3.0f - 2.0f * t);
  float width = 1.0f + (v.lod_blend_max - 1.0f) * smooth;
  bool blending = width > 1.001f;
  v.fade_scale = blending ? 255.0f / vg_logf(width) : 0.0f;
  float t0 = v.error_threshold * camera_scale;
  float t1 = (v.error_threshold * v.shadow_error_scale) * fmaxf(camera_scale, shadow_scale);
  v.low[0] = t0;
  v.high[0] = blending ? t0 * width : t0;
  v.low[1] = t1;
  v.high[1] = blending ? t1 * width : t1;
}


__device__ __forceinline__ uint32_t ceil_log2(uint32_t n) { return n <= 1 ? 0u : 32u - __clz(n - 1u); }

// The HZB level table row (offset, width, height, unused) of a level, read from the header: rows are 16 bytes
// and 16-byte aligned, so one load per lane.
struct HzbLevels {
  const float* __restrict__ hzb;
  __device__ __forceinline__ float4 operator()(uint32_t level) const { return reinterpret_cast<const float4*>(hzb)[level]; }
};

// The same rows from a wave's own copy of the table in LDS (loaded alongside the wave's other inputs, so the
// test needs one dependent global round trip, for the texels, instead of two).
struct HzbLevelsLds {
  const float4* table;
  __device__ __forceinline__ float4 operator()(uint32_t level) const { return table[level]; }
};

// occluded, OcclusionCulling.ts:119-170. `hzb` is only read when the sphere is in front of the near plane.
// The HZB header (sizes, level count) is read as wave-uniform data, the level rows and texels per lane.
template <typename V, typename Levels>
__device__ bool occluded_by(const V& v, const float* __restrict__ hzb, const Levels& levels, float vx, float vy, float vz,
                            float radius) {
  const auto header = scalar(hzb);
  float depth = -vz;
  float near_depth = depth - radius;
  float far_depth = depth + radius;
  if (near_depth <= v.near_plane) return false;
  const auto P = v.proj;
  float x_lo = vx - radius, x_hi = vx + radius, y_lo = vy - radius, y_hi = vy + radius;
  float sx = P[0], sy = P[5];
  float min_x, max_x, min_y, max_y;
  if (radius >= 0.0f && near_depth > 0.0f) {
    // 0 < near_depth <= far_depth: correctly rounded division is monotonic in the denominator,
    // so fminf(x / near_depth, x / far_depth) is x / far_depth for x >= 0 (signed zeros and NaN included)
    // and x / near_depth otherwise, and the reverse for fmaxf. One division per bound.
    min_x = (x_lo >= 0.0f ? x_lo / far_depth : x_lo / near_depth) * sx - P[8];
    max_x = (x_hi >= 0.0f ? x_hi / near_depth : x_hi / far_depth) * sx - P[8];
    min_y = (y_lo >= 0.0f ? y_lo / far_depth : y_lo / near_depth) * sy - P[9];
    max_y = (y_hi >= 0.0f ? y_hi / near_depth : y_hi / far_depth) * sy - P[9];
  } else {
    min_x = fminf(x_lo / near_depth, x_lo / far_depth) * sx - P[8];
    max_x = fmaxf(x_hi / near_depth, x_hi / far_depth) * sx - P[8];
    min_y = fminf(y_lo / near_depth, y_lo / far_depth) * sy - P[9];
    max_y = fmaxf(y_hi / near_depth, y_hi / far_depth) * sy - P[9];
  }
  float min_z = P[14] / near_depth - P[10];
  float W = header[1], H = header[2];
  float x0 = clampf(floorf((min_x * 0.5f + 0.5f) * W), 0.0f, W - 1.0f);
  float x1 = clampf(floorf((max_x * 0.5f + 0.5f) * W), 0.0f, W - 1.0f);
  float y0 = clampf(floorf((0.5f - max_y * 0.5f) * H), 0.0f, H - 1.0f);
  float y1 = clampf(floorf((0.5f - min_y * 0.5f) * H), 0.0f, H - 1.0f);
  uint32_t extent = static_cast<uint32_t>(fmaxf(x1 - x0, y1 - y0)) + 1u;
  uint32_t level = ceil_log2(extent);
  uint32_t top = static_cast<uint32_t>(header[kHzbLevelCountSlot]) - 1u;
  level = level > top ? top : level;
  // fx / 2^level == fx * 2^-level exactly (fx is a texel coordinate in [0, 4096): no underflow), so the
  // division by the mip scale is a multiplication by its exact reciprocal.
  const float inv_scale = __builtin_ldexpf(1.0f, -static_cast<int>(level));
  const float4 row = levels(level);  // level differs per lane
  const float offset = row.x, lw = row.y, lh = row.z;
  auto sample = [&](float fx, float fy) {
    float tx = clampf(floorf(fx * inv_scale), 0.0f, lw - 1.0f);
    float ty = clampf(floorf(fy * inv_scale), 0.0f, lh - 1.0f);
    return hzb[static_cast<uint32_t>(offset + ty * lw + tx)];
  };
  float m = fmaxf(fmaxf(sample(x0, y0), sample(x1, y0)), fmaxf(sample(x0, y1), sample(x1, y1)));
  return min_z < 1.0f && min_z > m + 1e-6f;
}

template <typename V>
__device__ __forceinline__ bool occluded(const V& v, const float* __restrict__ hzb, float vx, float vy, float vz, float radius) {
  return occluded_by(v, hzb, HzbLevels{hzb}, vx, vy, vz, radius);
}

// occludedSphere, GeometryPool.ts:701-705.
template <typename V, typename Levels>
__device__ __forceinline__ bool occluded_sphere_by(const Params& p, const V& v, const Levels& levels, float cx, float cy,
                                                   float cz, float radius) {
  if (!v.hzb_valid) return false;
  float4 vc = mat_mul16(v.view, cx, cy, cz, 1.0f);
  float margin = (fabsf(vc.z) * v.high[0]) / v.proj_scale;
  return occluded_by(v, scalar(p.hzbs)[v.slot], levels, vc.x, vc.y, vc.z, radius + margin);
}

template <typename V>
__device__ __forceinline__ bool occluded_sphere(const Params& p, const V& v, bool on, float cx, float cy, float cz,
                                                float radius) {
  return on && occluded_sphere_by(p, v, HzbLevels{scalar(p.hzbs)[v.slot]}, cx, cy, cz, radius);
}

// ------------------------------------------------------------------ 0. begin (resetNode, GeometryPool.ts:714-725)

__global__ __launch_bounds__(256) void k_begin(Params p) {
  ViewDev& v = p.views[blockIdx.x];
  if (threadIdx.x < 6) v.sweep_reach[threadIdx.x] = fmaxf(dot3(&v.planes[threadIdx.x * 4], v.sweep[0], v.sweep[1], v.sweep[2]), 0.0f);
  if (threadIdx.x == 6) view_bounds(v);
  if (threadIdx.x == 7) frustum_bounds(v);
  uint32_t* c = view_counters(p, v.slot);
  for (uint32_t i = threadIdx.x; i < kPerMeshCounters + 2 * p.slots * 4; i += blockDim.x) c[i] = 0;
  if (p.phase == 2) {
    // Second phase: the camera cut continues the cut's draw-list numbering, with its final band; there
    // is no shadow cut, retry or budget, and the slot's metrics add up over both phases.
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t* pw = phase_at(p, v.slot);
      *counter_at(p, v.slot, 0, kSelected) = p.counters1[static_cast<size_t>(v.slot) * p.counter_stride + 2 + kSelected];
      v.low[0] = __uint_as_float(pw[P_LOW]);
      v.high[0] = __uint_as_float(pw[P_HIGH]);
      v.fade_scale = __uint_as_float(pw[P_FADE]);
      v.low[1] = v.high[1] = 0.0f;
      v.redo[0] = v.redo[1] = 0;
      v.budget = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < kWorkParts) p.work_counter[threadIdx.x * kWorkCounterSpacing] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *p.cell_queue_count = 0;
      p.flags[0] = p.flags[1] = p.flags[2] = p.flags[3] = p.flags[4] = 0;
    }
    return;
  }
  if (threadIdx.x < kMetricWords) *metric_at(p, v.slot, threadIdx.x) = 0;
  if (v.two_phase && threadIdx.x < 2) phase_at(p, v.slot)[threadIdx.x] = 0;  // P_REJECTS, P_OVERFLOW
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t* state = state_at(p, v.slot);
    if (v.fresh)  // controller at rest, as GeometryPool.ts:602-603 creates it
      for (uint32_t i = 0; i < kStateWords; i++) state[i] = i == kStateScale || i == kStateScale + 1 ? 10000u : 0u;
    load_bands(v, state);
    v.redo[0] = v.redo[1] = 0;
    v.budget_bin = 0;
    *reinterpret_cast<float*>(metric_at(p, v.slot, M_THRESHOLD)) = v.error_threshold;
    *metric_at(p, v.slot, M_CELLS) = p.cell_count;  // the cell pass tests every cell in every view
    if (v.budget) {
      const float* t = p.budget_thresholds + static_cast<size_t>(blockIdx.x) * p.budget_bins;
      v.budget_scale = get_state(state, kStateScale);
      v.budget_low = t[0] * v.budget_scale;
      v.budget_high = t[p.budget_bins - 1] * v.budget_scale;
      const float ratio = t[p.budget_bins - 1] / t[0];
      v.budget_inv = p.budget_bins > 1 && ratio > 1.0f ? static_cast<float>(p.budget_bins - 1) / logf(ratio) : 0.0f;
    }
  }
  if (v.budget)
    for (uint32_t i = threadIdx.x; i <= p.budget_bins; i += blockDim.x)
      p.budget_hist[static_cast<size_t>(blockIdx.x) * (p.budget_bins + 1) + i] = 0;
  if (blockIdx.x == 0 && threadIdx.x < kWorkParts) p.work_counter[threadIdx.x * kWorkCounterSpacing] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *p.cell_queue_count = 0;
    p.flags[0] = p.flags[1] = p.flags[2] = p.flags[3] = p.flags[4] = 0;
  }
}

// ------------------------------------------------------------------ 1. cells (cellNode, GeometryPool.ts:728-760)

// One thread per cell, looping over a group of up to 64 launched views (blockIdx.y): the view records
// are wave-uniform (scalar loads), the cell's data is read once for all of them. The (view, cell) pairs
// that pass are appended cell-major with one atomic per wave, so the instance pass meets the views of
// one cell one after another and reuses the cell's instance data.
// A block takes 64 cells and up to 64 views: each of its 4 waves tests the 64 cells (one per lane) against 16
// of the views, the waves' view bits are OR-ed per cell in LDS, and the first wave appends the cells' visible
// (view, cell) pairs cell-major (views ascending within a cell) with one atomic, as one wave per 64 cells
// testing every view in turn would; the per-thread view loop is 4 times shorter.
__global__ __launch_bounds__(256) void k_cells(Params p) {
  __shared__ uint32_t visible[64];
  __shared__ unsigned long long cell_mask[5][64];  // kept views, then the views of each kQ* bit
  const uint32_t lane = lane_id(), group = threadIdx.x >> 6;
  const uint32_t cell = blockIdx.x * 64 + lane;
  const uint32_t v_first = blockIdx.y * 64, v_end = min(v_first + 64, p.view_count);
  const uint32_t g_first = v_first + group * 16, g_end = min(g_first + 16, v_end);
  if (threadIdx.x < 64) visible[threadIdx.x] = 0;
  for (uint32_t i = threadIdx.x; i < 5 * 64; i += blockDim.x) (&cell_mask[0][0])[i] = 0;
  __syncthreads();
  const bool valid = cell < p.cell_count;
  float4 sphere = make_float4(0, 0, 0, 0), record = sphere, flags = sphere, levers = sphere, bound = sphere;
  if (valid) {
    sphere = p.c_sphere[cell];
    record = p.c_rec[cell];
    bound = p.c_bound[cell];
    const uint32_t slot = static_cast<uint32_t>(record.y);
    flags = p.mesh_table[slot * kMeshStride + 5];
    levers = p.mesh_table[slot * kMeshStride + 3];
  }
  const bool enabled = valid && flags.x != 0.0f;
  const bool sweep = flags.y > 0.5f;
  const float radius = sphere.w, max_radius = record.x;
  uint64_t mask = 0, inside[4] = {0, 0, 0, 0};
  if (__ballot(enabled)) {
    for (uint32_t vi = g_first; vi < g_end; vi++) {
      const auto& v = scalar(p.views)[vi];
      bool keep = enabled && !(v.frustum == 1 && frustum_distance(v, sphere.x, sphere.y, sphere.z, 2, sweep) < -radius);
      const float min_pixels = fmaxf(v.min_pixel_radius, levers.y);
      const float4 vc = mat_mul16(v.view, sphere.x, sphere.y, sphere.z, 1.0f);
      const float view_distance = len3(vc.x, vc.y, vc.z);
      const float gap = view_distance - radius;
      const bool small = gap > max_radius && (max_radius * v.proj_scale) / fmaxf(gap, 1e-6f) < min_pixels;
      keep = keep && !small && !(gap > levers.x);
      // second phase: only cells holding an instance the cut rejected by occlusion
      if (p.phase == 2 && keep) keep = any_rejected(p, v.slot, static_cast<uint32_t>(record.z), min(static_cast<uint32_t>(record.w), kCellSize));
      if (keep) mask |= 1ull << (vi - v_first);
      // Frustum answers the whole cell settles for the instance pass (bounds on the cell's instances).
      if (v.frustum == 1 && __ballot(keep)) {
        const uint32_t bits = keep ? cell_inside_bits(v, sphere, bound, sweep) : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++)
          if (bits & (kQInCamera << k)) inside[k] |= 1ull << (vi - v_first);
      }
      const uint32_t n = static_cast<uint32_t>(__popcll(__ballot(keep)));
      if (lane == 0 && n) atomicAdd(&visible[vi - v_first], n);
    }
  }
  if (mask) atomicOr(&cell_mask[0][lane], mask);
#pragma unroll
  for (uint32_t k = 0; k < 4; k++)
    if (inside[k]) atomicOr(&cell_mask[1 + k][lane], inside[k]);
  __syncthreads();
  if (group == 0) {
    const uint64_t m_all = cell_mask[0][lane];
    const uint64_t m_in[4] = {cell_mask[1][lane], cell_mask[2][lane], cell_mask[3][lane], cell_mask[4][lane]};
    const uint32_t n = static_cast<uint32_t>(__popcll(m_all));
    const uint32_t incl = wave_scan(n);
    const uint32_t total = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
    uint32_t base = 0;
    if (lane == 0 && total) base = atomicAdd(p.cell_queue_count, total);
    uint32_t at = uniform(base) + incl - n;
    for (uint64_t m = m_all; m; m &= m - 1) {
      const uint32_t k = static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(m)) - 1);
      uint32_t bits = 0;
#pragma unroll
      for (uint32_t b = 0; b < 4; b++) bits |= ((m_in[b] >> k) & 1ull) ? (kQInCamera << b) : 0u;
      p.cell_queue[at++] = make_uint2((v_first + k) | bits, cell);
    }
  }
  if (threadIdx.x < v_end - v_first && visible[threadIdx.x]) {
    const uint32_t slot = scalar(p.views)[v_first + threadIdx.x].slot;
    atomicAdd(view_counters(p, slot), visible[threadIdx.x]);  // counters[0]: visible cells
  }
}

// ------------------------------------------------------------------ 2. instances (instancePass + selectLevels, GeometryPool.ts:771-913)

struct LevelResult {
  uint32_t first, last, possible, guaranteed;
  uint32_t levels;  // select_levels: the level giving `first` | the level giving `last` << 8
};

// selectLevels (GeometryPool.ts:771-800) for one band [low, high].
__device__ __forceinline__ LevelResult select_band(const Params& p, uint32_t level_base, uint32_t level_count, float low, float high,
                                                   float pmin, float pmax) {
  LevelResult r{0xffffffffu, 0u, 0u, 0u, 0u};
  for (uint32_t i = 0; i < level_count; i++) {
    const float4 l0 = p.mesh_table[level_base + i * 2];
    const float4 l1 = p.mesh_table[level_base + i * 2 + 1];
    const uint32_t level_first = static_cast<uint32_t>(l0.x);
    const uint32_t size = static_cast<uint32_t>(l0.y);
    const bool possible = (l0.z * pmax <= high) && (l0.w >= VG_INF_HALF || l0.w * pmin > low);
    const bool sure = (l1.x * pmin <= low * VG_MARGIN_LO) && (l1.y >= VG_INF_HALF || l1.y * pmax > high * VG_MARGIN_HI);
    if (possible) {
      r.first = min(r.first, level_first);
      r.last = max(r.last, level_first + size);
      r.possible += size;
      if (sure) r.guaranteed += size;
    }
  }
  return r;
}

// selectLevels for the camera and the shadow cut in one pass over the level table (the level records
// are wave-uniform loads): a level is possible if its error can project under `high` and its parent
// error over `low` somewhere in [pmin, pmax], guaranteed if that holds with margin at both ends.
// `levels`: the mesh's level table (2 float4 per level) staged in the wave's LDS; every lane reads the
// same entry (a broadcast read).
//
// Only levels in [first, last] are visited: outside it no lane's level can be possible (the caller's
// level_window).
template <bool kBoth, typename V>
__device__ __forceinline__ void select_levels(const float4* levels, const V& v, uint32_t first, uint32_t last, float pmin,
                                              float pmax, LevelResult (&r)[2]) {
  const float low0 = v.low[0], high0 = v.high[0], low1 = v.low[1], high1 = v.high[1];
  r[0] = r[1] = LevelResult{0xffffffffu, 0u, 0u, 0u, 0u};
  // first / last as min / max over the possible levels; `levels` records which level gave each (ties give
  // equal values, so either level's blocks are the same).
#pragma unroll 2
  for (uint32_t i = first; i <= last; i++) {
    const float4 l0 = levels[i * 2];
    const float4 l1 = levels[i * 2 + 1];
    const uint32_t level_first = static_cast<uint32_t>(l0.x);
    const uint32_t size = static_cast<uint32_t>(l0.y);
    const bool open_parent = l0.w >= VG_INF_HALF, open_guard = l1.y >= VG_INF_HALF;
    const float err_max = l0.z * pmax, parent_min = l0.w * pmin, guard_min = l1.x * pmin, guard_max = l1.y * pmax;
    {
      const bool possible = err_max <= high0 && (open_parent || parent_min > low0);
      const bool sure = guard_min <= low0 * VG_MARGIN_LO && (open_guard || guard_max > high0 * VG_MARGIN_HI);
      if (possible) {
        if (level_first < r[0].first) r[0].levels = (r[0].levels & ~255u) | i;
        if (level_first + size > r[0].last) r[0].levels = (r[0].levels & 255u) | (i << 8);
        r[0].first = min(r[0].first, level_first);
        r[0].last = max(r[0].last, level_first + size);
        r[0].possible += size;
        if (sure) r[0].guaranteed += size;
      }
    }
    if (kBoth) {
      const bool possible = err_max <= high1 && (open_parent || parent_min > low1);
      const bool sure = guard_min <= low1 * VG_MARGIN_LO && (open_guard || guard_max > high1 * VG_MARGIN_HI);
      if (possible) {
        if (level_first < r[1].first) r[1].levels = (r[1].levels & ~255u) | i;
        if (level_first + size > r[1].last) r[1].levels = (r[1].levels & 255u) | (i << 8);
        r[1].first = min(r[1].first, level_first);
        r[1].last = max(r[1].last, level_first + size);
        r[1].possible += size;
        if (sure) r[1].guaranteed += size;
      }
    }
  }
}

// The levels a wave's select_levels must visit, as [first, last] (first > last: none). Lane i tests level i
// against wave-wide bounds: a level is possible for a lane only if err * pmax <= high and parent * pmin > low;
// with pmax >= the wave's least pmax, pmin <= its largest pmin, high <= the larger band top and low >= the
// smaller band bottom, and binary32 products monotone in each factor, a level failing the test with those
// bounds fails it for every lane and both cuts. `active`: lanes whose instance is in a cut.
template <typename V>
__device__ __forceinline__ uint2 level_window(const float4* levels, const V& v, uint32_t level_count, bool any1, bool active,
                                              float pmin, float pmax) {
  // pmin, pmax > 0: their bit patterns order as the values do.
  const float least_pmax = __uint_as_float(~wave_max_u32(active ? ~__float_as_uint(pmax) : 0u));
  const float most_pmin = __uint_as_float(wave_max_u32(active ? __float_as_uint(pmin) : 0u));
  const float high = any1 ? fmaxf(v.high[0], v.high[1]) : v.high[0];
  const float low = any1 ? fminf(v.low[0], v.low[1]) : v.low[0];
  const uint32_t lane = lane_id();
  bool maybe = false;
  if (lane < level_count) {
    const float4 l0 = levels[lane * 2];
    maybe = l0.z * least_pmax <= high && (l0.w >= VG_INF_HALF || l0.w * most_pmin > low);
  }
  const uint64_t m = __ballot(maybe);
  if (!m) return make_uint2(1u, 0u);
  return make_uint2(static_cast<uint32_t>(__builtin_ctzll(m)), 63u - static_cast<uint32_t>(__builtin_clzll(m)));
}

// Sets bits [first, last) of a bitmap (measurement only).
__device__ void mark_range(uint32_t* bits, uint32_t first, uint32_t last) {
  for (uint32_t w = first >> 5; w <= (last - 1) >> 5; w++) {
    const uint32_t lo = w == (first >> 5) ? (first & 31) : 0u;
    const uint32_t hi = w == ((last - 1) >> 5) ? ((last - 1) & 31) : 31u;
    const uint32_t mask = (hi == 31 ? 0xffffffffu : ((1u << (hi + 1)) - 1u)) & ~((1u << lo) - 1u);
    atomicOr(&bits[w], mask);
  }
}

// Per-view counters of the instance pass, summed in a block's LDS when a launch has at most this many views
// (the queue is cell-major, so consecutive entries change view every time, while a block's entries touch
// few distinct views): flushed to memory once per block.
constexpr uint32_t kBlockViews = 64;
enum : uint32_t { A_INSTANCES, A_OCC_TESTS, A_OCCLUDED, A_SURVIVED0, A_SURVIVED1, A_WORK0, A_WORK1, A_CLUSTERS0, A_CLUSTERS1, kAccWords };

// The work-list partition of a queue entry's items (wave-uniform): by view when a launch has at least
// kViewPartsMin views, by wave rank otherwise (transient_sizes sizes the partitions for either).
__device__ __forceinline__ uint32_t work_part(const Params& p, uint32_t vi, uint64_t rank, uint64_t waves) {
  return p.view_count >= kViewPartsMin ? vi % kWorkParts : static_cast<uint32_t>(rank * kWorkParts / waves);
}

// One wave per queued (view, cell), the cell's 128 instances in two rounds of 64 lanes. Each
// (instance, cut) that survives the culls becomes at most one work item: a range of LOD runs for the
// meshlet pass, or, on the fast path, the instance's <= kFastPathMax guaranteed meshlets flagged to be
// drawn without cluster tests. The cell's items are appended with one atomic, lane-major, so item order
// and first-task order agree. Counters are summed in registers while the view stays the same, then in
// the block's LDS (or directly in memory when a launch has more than kBlockViews views).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_instances(Params p, uint32_t retry_round) {
  __shared__ float4 level_table[4][2 * kMaxLevels];  // per wave: the current mesh's level table
  __shared__ uint2 level_blk[4][kMaxLevels];         // per wave: its levels' first block, last block + 1
  __shared__ unsigned long long view_acc[kBlockViews][kAccWords];
  if (retry_round && !p.flags[retry_round - 1]) return;  // grid-uniform
  const bool block_acc = p.view_count <= kBlockViews;
  if (block_acc)
    for (uint32_t i = threadIdx.x; i < kBlockViews * kAccWords; i += blockDim.x) (&view_acc[0][0])[i] = 0ull;
  __syncthreads();
  const uint32_t lane = lane_id(), wib = threadIdx.x / 64;
  const uint32_t queue_count = *p.cell_queue_count;
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64), rank = xcd_wave_rank();
  const uint32_t q_begin = static_cast<uint32_t>(queue_count * rank / waves);
  const uint32_t q_end = static_cast<uint32_t>(queue_count * (rank + 1) / waves);
  if (q_begin < q_end) {
  // Counters of the current view (wave-uniform), flushed when the view changes.
  uint32_t acc_slot = 0xffffffffu, acc_vi = 0;
  uint32_t acc_instances = 0, acc_occ_tests = 0, acc_occluded = 0, acc_survived[2] = {0, 0}, acc_work[2] = {0, 0};
  uint64_t acc_clusters[2] = {0, 0};
  auto flush = [&]() {
    if (acc_slot == 0xffffffffu || lane != 0) return;
    if (block_acc) {
      unsigned long long* a = view_acc[acc_vi];
      const unsigned long long v[kAccWords] = {acc_instances, acc_occ_tests, acc_occluded, acc_survived[0], acc_survived[1], acc_work[0],
                                               acc_work[1], acc_clusters[0], acc_clusters[1]};
#pragma unroll
      for (uint32_t k = 0; k < kAccWords; k++)
        if (v[k]) atomicAdd(&a[k], v[k]);
      return;
    }
    if (acc_instances) atomicAdd(metric_at(p, acc_slot, M_INSTANCES), static_cast<unsigned long long>(acc_instances));
    if (acc_occ_tests) atomicAdd(metric_at(p, acc_slot, M_OCCLUSION), static_cast<unsigned long long>(acc_occ_tests));
    if (acc_occluded) atomicAdd(view_counters(p, acc_slot) + 1, acc_occluded);
    for (uint32_t c = 0; c < 2; c++) {
      if (acc_survived[c]) atomicAdd(counter_at(p, acc_slot, c, kSurvived), acc_survived[c]);
      if (acc_work[c]) {
        atomicAdd(counter_at(p, acc_slot, c, kWork), acc_work[c]);
        atomicAdd(metric_at(p, acc_slot, M_WORK0 + c), static_cast<unsigned long long>(acc_work[c]));
      }
      if (acc_clusters[c]) atomicAdd(metric_at(p, acc_slot, M_CLUSTERS0 + c), static_cast<unsigned long long>(acc_clusters[c]));
    }
  };

  // The current cell: its record, mesh and, per round of 64 instances, the view-independent values
  // (scale, bounding-sphere centre, LOD-sphere centre, culling-sphere centre), kept while consecutive
  // queue entries name the same cell.
  uint32_t cur_cell = 0xffffffffu, cur_mesh = 0xffffffffu, slot = 0, first_instance = 0, count = 0, level_count = 0, level_base = 0;
  float4 object_sphere = make_float4(0, 0, 0, 0), lod_bounds = object_sphere, cull_bounds = object_sphere;
  float4 flags = object_sphere, levers = object_sphere;
  float d_scale[2], d_center[2][3], d_lq[2][4], d_cc[2][3];
#pragma unroll
  for (uint32_t r = 0; r < 2; r++) {
    d_scale[r] = 0.0f;
    d_center[r][0] = d_center[r][1] = d_center[r][2] = 0.0f;
    d_lq[r][0] = d_lq[r][1] = d_lq[r][2] = d_lq[r][3] = 0.0f;
    d_cc[r][0] = d_cc[r][1] = d_cc[r][2] = 0.0f;
  }

  // The most items this wave's partition held after one of its appends (one atomicMax at the end).
  uint32_t high_water = 0;
  uint2 entry = scalar_u2(p.cell_queue, q_begin);
  for (uint32_t q = q_begin; q < q_end; q++) {
    const uint32_t vi = uniform(entry.x) & kQueueViewMask;
    const uint32_t qbits = uniform(entry.x) & ~kQueueViewMask;  // frustum answers settled for the whole cell
    const uint32_t cell = uniform(entry.y);
    // The work-list partition: with at least kViewPartsMin views, the entry's view's (a partition then holds the
    // items of few views, and the meshlet pass, which takes 64 consecutive tasks at a time and runs its
    // per-view loops once per view present, meets few views per chunk); else the wave's (ranks split evenly).
    const uint32_t part = work_part(p, vi, rank, waves);
    const size_t part_base = static_cast<size_t>(part) * p.work_capacity;
    if (q + 1 < q_end) entry = scalar_u2(p.cell_queue, q + 1);  // next entry, loaded while this one runs
    const auto& v = scalar(p.views)[vi];
    bool redo_camera = true, redo_shadow = p.phase != 2;  // the second phase has no shadow cut
    if (retry_round) {
      redo_camera = v.redo[0] != 0;
      redo_shadow = v.redo[1] != 0;
      if (!redo_camera && !redo_shadow) continue;
    }
    const bool record = p.phase == 0 && v.two_phase != 0 && redo_camera;
    const uint32_t vslot = uniform(v.slot);
    if (vslot != acc_slot) {
      flush();
      acc_slot = vslot;
      acc_vi = vi;
      acc_instances = acc_occ_tests = acc_occluded = 0;
      acc_survived[0] = acc_survived[1] = acc_work[0] = acc_work[1] = 0;
      acc_clusters[0] = acc_clusters[1] = 0;
    }
    if (cell != cur_cell) {
      cur_cell = cell;
      const float4 record = scalar_f4(p.c_rec, cell);
      slot = uniform(static_cast<uint32_t>(record.y));
      first_instance = uniform(static_cast<uint32_t>(record.z));
      count = min(uniform(static_cast<uint32_t>(record.w)), kCellSize);
      const float4* mesh = p.mesh_table + slot * kMeshStride;
      levers = scalar_f4(mesh, 3);
      level_count = min(static_cast<uint32_t>(levers.z), kMaxLevels);
      level_base = static_cast<uint32_t>(levers.w);
      object_sphere = scalar_f4(mesh, 0);
      lod_bounds = scalar_f4(mesh, 1);
      cull_bounds = scalar_f4(mesh, 2);
      flags = scalar_f4(mesh, 5);
      if (slot != cur_mesh) {  // stage the mesh's level table in LDS (one coalesced load)
        cur_mesh = slot;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < 2 * level_count; i += 64) level_table[wib][i] = p.mesh_table[level_base + i];
        for (uint32_t i = lane; i < level_count; i += 64) level_blk[wib][i] = p.level_blocks[(level_base - p.slots * kMeshStride) / 2 + i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      float4 m[2][4];
#pragma unroll
      for (uint32_t r = 0; r < 2; r++) {
        const uint32_t instance = first_instance + r * 64 + lane;
        const bool valid = r * 64 + lane < count;
        m[r][0] = valid ? p.i_col0[instance] : make_float4(0, 0, 0, 0);
        m[r][1] = valid ? p.i_col1[instance] : make_float4(0, 0, 0, 0);
        m[r][2] = valid ? p.i_col2[instance] : make_float4(0, 0, 0, 0);
        m[r][3] = valid ? p.i_col3[instance] : make_float4(0, 0, 0, 1);
      }
#pragma unroll
      for (uint32_t r = 0; r < 2; r++) {
        const float4 c0 = m[r][0], c1 = m[r][1], c2 = m[r][2], c3 = m[r][3];
        d_scale[r] = fmaxf(len3(c0.x, c0.y, c0.z), fmaxf(len3(c1.x, c1.y, c1.z), len3(c2.x, c2.y, c2.z)));
        const float4 center = mat_mul(c0, c1, c2, c3, object_sphere.x, object_sphere.y, object_sphere.z, 1.0f);
        d_center[r][0] = center.x;
        d_center[r][1] = center.y;
        d_center[r][2] = center.z;
        const float4 lq = mat_mul(c0, c1, c2, c3, lod_bounds.x, lod_bounds.y, lod_bounds.z, 1.0f);
        d_lq[r][0] = lq.x;
        d_lq[r][1] = lq.y;
        d_lq[r][2] = lq.z;
        d_lq[r][3] = lq.w;
        const float4 cc = mat_mul(c0, c1, c2, c3, cull_bounds.x, cull_bounds.y, cull_bounds.z, 1.0f);
        d_cc[r][0] = cc.x;
        d_cc[r][1] = cc.y;
        d_cc[r][2] = cc.z;
      }
    }
    const bool shadow_on = flags.y > 0.5f, occlusion_on = flags.z > 0.5f;
    const float min_pixels = fmaxf(v.min_pixel_radius, levers.y);
    const float max_draw_distance = levers.x;
    const bool occlusion_test = occlusion_on && v.hzb_valid != 0;
    acc_instances += count;

    // The lane's items: (round, cut) -> task range [lo, hi) (empty when lo == hi), fast flags.
    uint32_t item_lo[2][2], item_hi[2][2], fast_bits = 0;
    float item_scale[2];
    uint32_t lane_clusters[2] = {0, 0};
#pragma unroll
    for (uint32_t r = 0; r < 2; r++) {
      item_lo[r][0] = item_hi[r][0] = item_lo[r][1] = item_hi[r][1] = 0;
      item_scale[r] = 0.0f;
      if (r * 64 >= count) continue;  // wave-uniform
      const uint32_t instance = first_instance + r * 64 + lane;
      bool valid = r * 64 + lane < count;
      if (p.phase == 2 && valid)  // the second phase re-tests the instances the cut rejected by occlusion
        valid = (p.reject_bits[static_cast<size_t>(vslot) * p.reject_words + (instance >> 5)] >> (instance & 31u)) & 1u;
      const float scale = d_scale[r];
      item_scale[r] = scale;
      const float3 center = make_float3(d_center[r][0], d_center[r][1], d_center[r][2]);
      const float radius = object_sphere.w * scale;
      const bool no_frustum = v.frustum == 0;
      bool in_camera = valid && redo_camera &&
                       (no_frustum || (qbits & kQInCamera) || frustum_distance(v, center.x, center.y, center.z, 0, false) >= -radius);
      const bool in_shadow = valid && redo_shadow && shadow_on &&
                             (no_frustum || (qbits & kQInShadow) || frustum_distance(v, center.x, center.y, center.z, 1, false) >= -radius);
      const float4 vc = mat_mul16(v.view, center.x, center.y, center.z, 1.0f);
      const float view_distance = len3(vc.x, vc.y, vc.z);
      const bool too_small = view_distance > radius && (radius * v.proj_scale) / view_distance < min_pixels;
      const bool too_far = (view_distance - radius) > max_draw_distance;
      const bool keep = (in_camera || in_shadow) && !too_small && !too_far;
      if (p.fp_instances && valid) atomicOr(&p.fp_instances[instance >> 5], 1u << (instance & 31));
      const bool test = keep && in_camera && occlusion_test;
      bool occ = false;
      if (test) occ = occluded_sphere(p, v, true, center.x, center.y, center.z, radius);
      acc_occ_tests += static_cast<uint32_t>(__popcll(__ballot(test)));
      acc_occluded += static_cast<uint32_t>(__popcll(__ballot(occ)));
      in_camera = in_camera && !occ;
      if (record) {
        const uint64_t occluded = __ballot(occ);
        if (lane == 0) record_instance_rejects(p, vslot, first_instance + r * 64, min(64u, count - r * 64), occluded);
      }

      const bool in_cut[2] = {keep && in_camera, keep && in_shadow};
      const bool any0 = __ballot(in_cut[0]) != 0, any1 = __ballot(in_cut[1]) != 0;
      if (!any0 && !any1) continue;
      const float4 lv = mat_mul16(v.view, d_lq[r][0], d_lq[r][1], d_lq[r][2], d_lq[r][3]);
      const float d = len3(lv.x, lv.y, lv.z);
      const float rr = lod_bounds.w * scale;
      const float pmin = (scale * v.proj_scale) / fmaxf(d - rr, v.near_plane);
      const float pmax = (scale * v.proj_scale) / fmaxf(d + rr, v.near_plane);
      LevelResult lr[2];
      const uint2 window = level_window(level_table[wib], v, level_count, any1, in_cut[0] || in_cut[1], pmin, pmax);
      if (any1) select_levels<true>(level_table[wib], v, window.x, window.y, pmin, pmax, lr);
      else select_levels<false>(level_table[wib], v, window.x, window.y, pmin, pmax, lr);
      const float3 cc = make_float3(d_cc[r][0], d_cc[r][1], d_cc[r][2]);
      const float cull_radius = cull_bounds.w * scale;
      bool work[2];
#pragma unroll
      for (uint32_t cut = 0; cut < 2; cut++) {
        const bool survived = in_cut[cut] && lr[cut].last > lr[cut].first;
        const bool fast = survived && lr[cut].possible == lr[cut].last - lr[cut].first && lr[cut].guaranteed == lr[cut].possible &&
                          lr[cut].possible <= kFastPathMax &&
                          (no_frustum || (qbits & (cut == 1 ? kQCullShadow : kQCullCamera)) ||
                           frustum_distance(v, cc.x, cc.y, cc.z, cut == 1 ? 1 : 0, false) >= cull_radius);
        work[cut] = survived && !fast;
        acc_survived[cut] += static_cast<uint32_t>(__popcll(__ballot(survived)));
        acc_work[cut] += static_cast<uint32_t>(__popcll(__ballot(work[cut])));
        if (survived) lane_clusters[cut] += lr[cut].last - lr[cut].first;
        if (p.fp_meshlets && survived) mark_range(p.fp_meshlets, lr[cut].first, lr[cut].last);
        if (fast) {
          item_lo[r][cut] = lr[cut].first;
          item_hi[r][cut] = lr[cut].last;
          fast_bits |= 1u << (r * 2 + cut);
        }
      }
      // Work items cover the LOD blocks of [first, last) (level ranges hold whole runs, runs and blocks
      // never cross a level's bounds): from the block of the first level's first meshlet to that of the last
      // level's last one.
#pragma unroll
      for (uint32_t cut = 0; cut < 2; cut++) {
        if (work[cut]) {
          item_lo[r][cut] = level_blk[wib][lr[cut].levels & 255u].x;
          item_hi[r][cut] = level_blk[wib][lr[cut].levels >> 8].y;
        }
      }
    }
    acc_clusters[0] += wave_sum(lane_clusters[0]);
    acc_clusters[1] += wave_sum(lane_clusters[1]);

    // Append the cell's items: lane-major, (round, cut) order within a lane; one atomic on the wave's
    // partition counter places them and gives their first task.
    uint32_t n_items = 0, n_tasks = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      const uint32_t len = item_hi[k >> 1][k & 1] - item_lo[k >> 1][k & 1];
      n_items += len ? 1u : 0u;
      n_tasks += len;
    }
    const uint32_t items_incl = wave_scan(n_items);
    const uint32_t items = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(items_incl), 63));
    if (items == 0) continue;
    const uint32_t tasks_incl = wave_scan(n_tasks);
    const uint32_t tasks = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(tasks_incl), 63));
    unsigned long long old = 0;
    if (lane == 0) old = atomicAdd(p.work_counter + part * kWorkCounterSpacing, (static_cast<unsigned long long>(items) << kTaskBits) | tasks);
    const uint64_t o = (static_cast<uint64_t>(uniform(static_cast<uint32_t>(old >> 32))) << 32) | uniform(static_cast<uint32_t>(old));
    const uint32_t item_base = static_cast<uint32_t>(o >> kTaskBits);
    const uint64_t task_base = o & kTaskMask;
    if (lane == 0 && (task_base + tasks >= kTaskLimit || static_cast<uint64_t>(item_base) + items > p.work_capacity))
      atomicOr(&p.flags[3], 1u);
    high_water = max(high_water, item_base + items);
    uint32_t at = item_base + items_incl - n_items;
    uint64_t task_at = task_base + tasks_incl - n_tasks;
    const uint32_t tag_base = slot | (vi << kTagViewShift);
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      const uint32_t r = k >> 1, cut = k & 1;
      const uint32_t lo = item_lo[r][cut], hi = item_hi[r][cut];
      if (hi > lo) {
        if (at < p.work_capacity) {
          const uint32_t tag = tag_base | (cut ? kTagShadow : 0u) | (cut == 0 && occlusion_on ? kTagOcclusion : 0u) |
                               ((fast_bits >> k) & 1u ? kTagFast : 0u);
          const size_t g = part_base + at;
          p.work[g] = make_uint4(first_instance + r * 64 + lane, lo, hi, tag);
          p.work_start[g] = task_at;
          p.work_scale[g] = item_scale[r];
        }
        at++;
        task_at += hi - lo;
      }
    }
  }
  flush();
  if (lane == 0 && high_water) atomicMax(&p.flags[4], high_water);
  }
  if (!block_acc) return;  // block-uniform
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < p.view_count * kAccWords; i += blockDim.x) {
    const unsigned long long wide = (&view_acc[0][0])[i];
    if (!wide) continue;
    const uint32_t value = static_cast<uint32_t>(wide);  // the 32-bit counters' words (their sums fit 32 bits)
    const uint32_t vi = i / kAccWords, k = i % kAccWords, slot = p.views[vi].slot;
    switch (k) {
      case A_INSTANCES: atomicAdd(metric_at(p, slot, M_INSTANCES), wide); break;
      case A_OCC_TESTS: atomicAdd(metric_at(p, slot, M_OCCLUSION), wide); break;
      case A_OCCLUDED: atomicAdd(view_counters(p, slot) + 1, value); break;
      case A_SURVIVED0: atomicAdd(counter_at(p, slot, 0, kSurvived), value); break;
      case A_SURVIVED1: atomicAdd(counter_at(p, slot, 1, kSurvived), value); break;
      case A_WORK0:
      case A_WORK1:
        atomicAdd(counter_at(p, slot, k - A_WORK0, kWork), value);
        atomicAdd(metric_at(p, slot, M_WORK0 + (k - A_WORK0)), wide);
        break;
      default: atomicAdd(metric_at(p, slot, M_CLUSTERS0 + (k - A_CLUSTERS0)), wide); break;
    }
  }
}

// ------------------------------------------------------------------ 3. meshlets (meshletPass, GeometryPool.ts:926-968)

// Walks the local tasks [lb, le) of one work-list partition (its items at [base, base + items)), 64
// consecutive tasks at a time: the wave finds the item holding lb by a 64-ary search over the items'
// first tasks, then maps lanes to items with a 6-step shuffle search over the next 64 item starts, whose
// load for the next chunk is issued before the current chunk's body runs.
template <typename Body>
__device__ __forceinline__ void walk_partition(const Params& p, size_t base, uint32_t items, uint64_t lb, uint64_t le, Body&& body) {
  const uint32_t lane = lane_id();
  const unsigned long long* __restrict__ starts = p.work_start + base;
  const uint4* __restrict__ work = p.work + base;
  // The item holding lb: the last item whose first task is <= lb (64-ary search).
  uint32_t lo = 0, hi = items;
  while (hi - lo > 1) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint32_t probe = lo + lane * step;
    const bool le_ = probe < hi && starts[probe] <= lb;
    uint32_t n = static_cast<uint32_t>(__popcll(__ballot(le_)));  // >= 1: starts[lo] <= lb
    n = n < 1 ? 1u : n;
    const uint32_t nlo = lo + (n - 1) * step;
    hi = min(hi, nlo + step);
    lo = nlo;
  }
  uint32_t cur = uniform(lo);
  uint64_t cur_start = starts[cur];
  uint64_t ns = cur + 1 + lane < items ? starts[cur + 1 + lane] : ~0ull;

  for (uint64_t k0 = lb; k0 < le; k0 += 64) {
    // Distances from k0 to the first tasks of the next 64 items (at least 1, clamped to 65). Item starts
    // strictly increase, so the items starting inside this chunk are the set bits of `starts_here`
    // (bit t: an item starts at task k0 + t): a lane's item is `cur` plus the number of starts at or
    // before it, and the item's first task is the highest such start.
    const uint32_t dist = ns - k0 > 65 ? 65u : static_cast<uint32_t>(ns - k0);
    const uint64_t bit = dist < 64 ? 1ull << dist : 0ull;
    const uint64_t starts_here = (static_cast<uint64_t>(wave_or(static_cast<uint32_t>(bit >> 32))) << 32) |
                                 wave_or(static_cast<uint32_t>(bit));
    const uint64_t upto = starts_here & ((2ull << lane) - 1ull);  // starts at tasks k0 + 1 .. k0 + lane
    const uint32_t pos = static_cast<uint32_t>(__popcll(upto));
    const uint64_t k = k0 + lane;
    const uint32_t item = cur + pos;
    const uint64_t item_start = pos == 0 ? cur_start : k0 + (63 - static_cast<uint32_t>(__clzll(upto)));
    bool valid = k < le && item < items && k >= item_start;

    // The item holding task k0 + 64, and the first tasks of its successors, loaded while this chunk runs.
    const uint32_t last_pos = static_cast<uint32_t>(__popcll(starts_here));
    uint32_t next_cur = cur + last_pos;
    uint64_t next_start = last_pos ? k0 + (63 - static_cast<uint32_t>(__clzll(starts_here))) : cur_start;
    if (__ballot(dist == 64)) {  // an item starts exactly at the next chunk
      next_cur += 1;
      next_start = k0 + 64;
    }
    next_cur = uniform(next_cur);
    if (k0 + 64 < le) ns = next_cur + 1 + lane < items ? ld32(starts, next_cur + 1 + lane) : ~0ull;

    uint4 w = make_uint4(0, 0, 0, 0);
    if (valid) w = ld32(work, item);
    const uint32_t task = w.y + static_cast<uint32_t>(k - item_start);
    if (valid && task >= w.z) {  // the walk lost its item: report instead of reading past the range
      atomicOr(&p.flags[2], 1u);
      valid = false;
      w = make_uint4(0, 0, 0, 0);
    }
    body(valid, static_cast<uint32_t>(base + item), w, task);
    cur = next_cur;
    cur_start = next_start;
  }
}

// Calls body(valid, item, w, task) for 64 consecutive tasks at a time over this wave's equal share of
// the flat task space: the partitions' task ranges, one after another (all lanes together).
template <typename Body>
__device__ __forceinline__ void for_each_task(const Params& p, Body&& body) {
  if (p.flags[3]) return;
  const uint32_t lane = lane_id();
  const unsigned long long packed = lane < kWorkParts ? p.work_counter[lane * kWorkCounterSpacing] : 0ull;
  const uint32_t part_items = min(static_cast<uint32_t>(packed >> kTaskBits), p.work_capacity);
  const uint64_t part_tasks = part_items ? (packed & kTaskMask) : 0ull;
  const uint64_t part_end = wave_scan64(part_tasks);  // end of each partition's tasks in the flat space
  const uint64_t total = readlane64(part_end, 63);
  if (total == 0) return;
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
  const uint64_t wave = xcd_wave_rank();
  const uint64_t seg_begin = total * wave / waves;
  const uint64_t seg_end = total * (wave + 1) / waves;
  for (uint64_t k0 = seg_begin; k0 < seg_end;) {
    // The partition holding task k0: the first whose end is past it.
    const uint32_t q = static_cast<uint32_t>(__popcll(__ballot(lane < kWorkParts && part_end <= k0)));
    if (q >= kWorkParts) break;  // cannot happen: k0 < total
    const uint64_t q_end = readlane64(part_end, q);
    const uint64_t q_start = q_end - readlane64(part_tasks, q);
    const uint64_t stop = min(seg_end, q_end);
    walk_partition(p, static_cast<size_t>(q) * p.work_capacity,
                   static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(part_items), static_cast<int>(q))), k0 - q_start,
                   stop - q_start, body);
    k0 = stop;
  }
}

// The meshlet pass in three stages per wave. (1) Blocks: a lane per block of the work items' ranges; the
// block test (above) drops blocks none of whose runs can be in the band; blocks that remain go to a
// per-wave LDS queue. (2) Runs: whenever 64 runs are queued, a lane per run runs the exact LOD test (own
// <= high, parent > low, the blend range), once per LOD run (consecutive meshlets whose LOD test has
// identical inputs); the meshlets of runs in the cut go to a per-wave LDS ring. (3) Visibility: full waves
// of those meshlets get the frustum and occlusion tests and are compacted (drawn = visible and in the cut,
// so the order changes no decision). Fast-path items (guaranteed meshlets) go straight to the ring.
// Compiled for several occupancy targets (waves per SIMD); the launch schedule picks one. kAffine: every
// instance matrix of the scene is affine and every run's LOD spheres finite (projected_error); only then
// are blocks dropped (the bounds assume the transformed w is 1).
template <int kWavesPerSimd, bool kAffine>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd))) void k_meshlets(Params p) {
  // Per wave: meshlets in the cut awaiting their visibility test, 16 bytes each: (instance | fade_low << 24,
  // meshlet | fade_high << 24, tag, instance scale bits). A scene holds fewer than 2^24 instances and meshlets.
  __shared__ uint4 ring[4][kQueue];
  // Per wave: the HZB level table of the view whose meshlets are being tested (one float per lane to copy).
  static_assert(kHzbMaxLevels * 4 == 64, "the level table is copied one float per lane");
  __shared__ float4 hzb_rows[4][kHzbMaxLevels];
  // Per wave: blocks that passed, (instance | run count << 24, first run, tag, instance scale bits), and the
  // wave-relative position of each one's first run in the stream of queued runs.
  static_assert((kBlockQueue & (kBlockQueue - 1)) == 0 && kBlockRuns < 256, "block queue indices wrap by masking");
  __shared__ uint4 bq[4][kBlockQueue];
  __shared__ uint32_t bq_start[4][kBlockQueue];
  const uint32_t wib = threadIdx.x / 64, lane = lane_id();
  uint32_t head = 0, tail = 0;  // wave-uniform ring positions
  // Block queue: entries [bq_head, bq_tail) (bq_head may be a consumed one), runs [rq_pos, rq_end) untested.
  // It starts with an empty consumed entry at position 0, so later entries always follow the head.
  uint32_t bq_head = 0, bq_tail = 1, rq_pos = 0, rq_end = 0;
  if (lane == 0) {
    bq[wib][0] = make_uint4(0, 0, 0, 0);
    bq_start[wib][0] = 0;
  }
  // Occlusion tests of the current view slot, added to its metric when the slot changes (and at the end).
  uint32_t occ_slot = 0xffffffffu, occ_count = 0;
  auto occ_flush = [&]() __attribute__((always_inline)) {
    if (occ_count && lane == 0) atomicAdd(metric_at(p, occ_slot, M_OCCLUSION), static_cast<unsigned long long>(occ_count));
    occ_count = 0;
  };

  auto wave_sync = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  auto drain = [&](uint32_t n) __attribute__((always_inline)) {
    const bool active = lane < n;
    const uint4 c = active ? ring[wib][(head + lane) & (kQueue - 1)] : make_uint4(0, 0, 0, 0);
    const float scale = __uint_as_float(c.w);
    const uint32_t instance = c.x & 0xffffffu, meshlet = c.y & 0xffffffu, tag = c.z;
    const uint32_t fades = (c.x >> 24) | ((c.y >> 24) << 8);
    const uint32_t mslot = tag & kTagSlotMask, cut = (tag & kTagShadow) ? 1u : 0u, vi = tag >> kTagViewShift;
    const bool occ_on = (tag & kTagOcclusion) != 0, fast = (tag & kTagFast) != 0;
    const uint32_t meshlet_tris = active ? ld32(p.m_tris, meshlet) : 0u;
    bool drawn = active && fast;  // fast path: guaranteed meshlets of a culling sphere inside the frustum
    bool occ_test = false, rejected = false;
    uint32_t vslot = 0;
    // The instance columns and cull sphere of every tested lane, loaded before the per-view loop so that
    // their latencies overlap (and are paid once however many views the wave holds).
    const bool tested = active && !fast;
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0, c3 = c0, cs = c0;
    if (tested) {
      c0 = ld32(p.i_col0, instance);
      c1 = ld32(p.i_col1, instance);
      c2 = ld32(p.i_col2, instance);
      c3 = ld32(p.i_col3, instance);
      cs = ld32(p.m_cull, meshlet);
    }
    // Visibility, one launched view at a time (its constants are wave-uniform there).
    for (uint64_t pending = __ballot(active); pending;) {
      const uint32_t vu = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vi), __ffsll(static_cast<unsigned long long>(pending)) - 1));
      const bool mine = active && vi == vu;
      // readfirstlane keeps the index uniform even where the compiler substitutes vi for vu.
      const auto& v = scalar(p.views)[uniform(vu)];
      if (v.hzb_valid && __ballot(mine && tested && cut == 0 && occ_on)) {
        // Copy the view's HZB level table to the wave's LDS rows (all lanes: the table is 64 floats).
        reinterpret_cast<float*>(hzb_rows[wib])[lane] = scalar(p.hzbs)[v.slot][lane];
        wave_sync();
      }
      if (mine) {
        vslot = v.slot;
        if (!fast) {
          const float4 center = mat_mul(c0, c1, c2, c3, cs.x, cs.y, cs.z, 1.0f);
          bool visible = v.frustum == 0 || frustum_distance(v, center.x, center.y, center.z, cut == 1 ? 1 : 0, false) >= -(cs.w * scale);
          occ_test = cut == 0 && visible && occ_on && v.hzb_valid;
          if (occ_test) visible = !occluded_sphere_by(p, v, HzbLevelsLds{hzb_rows[wib]}, center.x, center.y, center.z, cs.w * scale);
          rejected = occ_test && !visible && v.two_phase && p.phase == 0;
          drawn = visible;
        }
      }
      pending &= ~__ballot(mine);
    }
    const uint32_t tris = drawn ? meshlet_tris : 0u;
    head += n;
    const uint32_t packed_tris = tris | (fades << 8);
    for (uint64_t tests = __ballot(occ_test); tests;) {
      const uint32_t key = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vslot), __ffsll(static_cast<unsigned long long>(tests)) - 1));
      const uint64_t same = __ballot(occ_test && vslot == key);
      if (key != occ_slot) {
        occ_flush();
        occ_slot = key;
      }
      occ_count += static_cast<uint32_t>(__popcll(same));
      tests &= ~same;
    }
    if (__ballot(rejected)) {  // two-phase views: keep the occluded meshlets for the second phase
      const uint32_t rat = wave_atomic_add<true>(phase_at(p, vslot) + P_REJECTS, 1u, rejected);
      if (rejected && rat < p.reject_capacity)
        p.reject_meshlets[static_cast<size_t>(vslot) * p.reject_capacity + rat] =
            make_uint4(instance, meshlet, fades | (mslot << 16), __float_as_uint(scale));
      else if (rejected)
        atomicOr(phase_at(p, vslot) + P_OVERFLOW, 1u);
    }
    const uint32_t at = wave_atomic_add<true>(counter_at(p, vslot, cut, kSelected), 1u, drawn);
    const bool write = drawn && at < p.max_meshlets;
    const uint32_t part = (packed_tris & 0xffff00u) == kFullFadeBits ? 0u : 1u;
    const uint32_t first_index = wave_atomic_add(mesh_count_at(p, vslot, cut, mslot, part), tris * 3, write);
    if (write) p.draws[vslot][at + cut * p.max_meshlets] = make_uint4(instance, meshlet, first_index, packed_tris);
  };

  // Appends `count` (<= kMaxRun) meshlets from `first` per lane to the ring, the lanes' k-th meshlets
  // together for k = 0, 1, ..., and drains a full wave whenever 64 wait (so at most 63 + 64 are queued).
  // Ring positions run freely; their difference is what counts.
  auto ring_append = [&](uint32_t count, uint32_t first, uint32_t fades, uint32_t instance, uint32_t tag, float scale) __attribute__((always_inline)) {
    for (uint32_t k = 0; k < kMaxRun; k++) {
      const bool has = k < count;
      const uint64_t round = __ballot(has);
      if (!round) break;
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(round >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(round), 0u));
      if (has)
        ring[wib][(tail + below) & (kQueue - 1)] =
            make_uint4(instance | (fades << 24), (first + k) | ((fades >> 8) << 24), tag, __float_as_uint(scale));
      tail += static_cast<uint32_t>(__popcll(round));
      if (tail - head >= 64) {
        wave_sync();
        drain(64);
        wave_sync();
      }
    }
  };

  // Tests the next n (<= 64) queued runs, one per lane: the exact LOD test, then the ring.
  auto run_stage = [&](uint32_t n) __attribute__((always_inline)) {
    wave_sync();
    // The entries after the head that start inside [rq_pos, rq_pos + 63]: bit d of `starts` = one starts at
    // rq_pos + d. Each entry has at least one run, so no more than 64 of them start inside the window.
    const uint32_t e_next = bq_head + 1 + lane;
    const bool present = lane < bq_tail - bq_head - 1;
    const uint32_t dist = present ? bq_start[wib][e_next & (kBlockQueue - 1)] - rq_pos : 0xffffffffu;
    const uint64_t bit = dist < 64 ? 1ull << dist : 0ull;
    const uint64_t starts = (static_cast<uint64_t>(wave_or(static_cast<uint32_t>(bit >> 32))) << 32) | wave_or(static_cast<uint32_t>(bit));
    const bool valid = lane < n;
    const uint32_t e = bq_head + static_cast<uint32_t>(__popcll(starts & ((2ull << lane) - 1ull)));
    uint4 q = make_uint4(0, 0, 0, 0);
    uint32_t first_pos = 0;
    if (valid) {
      q = bq[wib][e & (kBlockQueue - 1)];
      first_pos = bq_start[wib][e & (kBlockQueue - 1)];
    }
    const uint32_t instance = q.x & 0xffffffu, tag = q.z, run = q.y + (rq_pos + lane - first_pos);
    const float scale = __uint_as_float(q.w);
    const uint32_t cut = (tag & kTagShadow) ? 1u : 0u, vi = tag >> kTagViewShift;
    // The run's records and its instance's columns, loaded before the per-view loop so that their latencies
    // overlap (the span too: loading it only for runs in the cut would add a dependent round trip).
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0, c3 = c0, lod_sphere = c0, parent_sphere = c0;
    float2 err = make_float2(0.0f, 0.0f);
    uint2 span = make_uint2(0, 0);
    if (valid) {
      c0 = ld32(p.i_col0, instance);
      c1 = ld32(p.i_col1, instance);
      c2 = ld32(p.i_col2, instance);
      c3 = ld32(p.i_col3, instance);
      err = ld32(p.r_err, run);
      lod_sphere = ld32(p.r_lod, run);
      parent_sphere = ld32(p.r_parent, run);
      span = ld32(p.r_span, run);
    }
    uint32_t fades = 0, count = 0;
    // The LOD test of a run, one launched view at a time (its constants are wave-uniform there).
    for (uint64_t pending = __ballot(valid); pending;) {
      const uint32_t vu = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vi), __ffsll(static_cast<unsigned long long>(pending)) - 1));
      const bool mine = valid && vi == vu;
      if (mine) {
        // readfirstlane keeps the index uniform even where the compiler substitutes vi for vu.
        const auto& v = scalar(p.views)[uniform(vu)];
        const float own = projected_error<kAffine>(v, c0, c1, c2, c3, scale, lod_sphere, err.x);
        const float parent = projected_error<kAffine>(v, c0, c1, c2, c3, scale, parent_sphere, err.y);
        const float low = cut ? v.low[1] : v.low[0], high = cut ? v.high[1] : v.high[0], fade = v.fade_scale;
        uint32_t fade_low = 0, fade_high = kFadeSteps;
        if (fade > 0.0f) {
          fade_low = static_cast<uint32_t>(clampf(rintf(vg_logf(fmaxf(own, 1e-30f) / low) * fade), 0.0f, 255.0f));
          fade_high = static_cast<uint32_t>(clampf(rintf(vg_logf(fmaxf(parent, 1e-30f) / low) * fade), 0.0f, 255.0f));
        }
        const bool lod = own <= high && parent > low && fade_high > fade_low;
        fades = fade_low | (fade_high << 8);
        if (lod) count = span.y;
      }
      pending &= ~__ballot(mine);
    }
    // Entries whose runs are all tested now: the new head is the last one starting at or before rq_pos + n.
    const uint64_t upto_n = n >= 64 ? ~0ull : ((2ull << n) - 1ull);
    const bool at_end = __ballot(dist == 64) != 0 && n >= 64;
    bq_head += static_cast<uint32_t>(__popcll(starts & upto_n)) + (at_end ? 1u : 0u);
    rq_pos += n;
    ring_append(count, span.x, fades, instance, tag, scale);
  };

  for_each_task(p, [&](bool valid, uint32_t item, const uint4& w, uint32_t task) __attribute__((always_inline)) {
    const bool fast = valid && (w.w & kTagFast) != 0;
    const bool test = valid && !fast;
    // Fast-path items: the task is a meshlet, drawn at full fade.
    if (__ballot(fast)) ring_append(fast ? 1u : 0u, task, kFadeSteps << 8, w.x, w.w, fast ? p.work_scale[item] : 0.0f);
    if (!__ballot(test)) return;
    // Blocks: the block's span and bounds and the instance's columns, loaded before the per-view loop.
    uint2 bspan = make_uint2(0, 0);
    float scale = 0.0f;
    if (test) {
      bspan = ld32(p.b_span, task);
      scale = p.work_scale[item];
    }
    bool keep = test;
    if (kAffine) {
      float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0, c3 = c0, blod = c0, bpar = c0, berr = c0;
      if (test) {
        c0 = ld32(p.i_col0, w.x);
        c1 = ld32(p.i_col1, w.x);
        c2 = ld32(p.i_col2, w.x);
        c3 = ld32(p.i_col3, w.x);
        blod = ld32(p.b_lod, task);
        bpar = ld32(p.b_par, task);
        berr = ld32(p.b_err, task);
      }
      const BlockBounds ib = instance_bounds(c0, c1, c2, c3, scale);
      const uint32_t cut = (w.w & kTagShadow) ? 1u : 0u, vi = w.w >> kTagViewShift;
      for (uint64_t pending = __ballot(test); pending;) {
        const uint32_t vu = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vi), __ffsll(static_cast<unsigned long long>(pending)) - 1));
        const bool mine = test && vi == vu;
        if (mine) {
          const auto& v = scalar(p.views)[uniform(vu)];
          const float low = cut ? v.low[1] : v.low[0], high = cut ? v.high[1] : v.high[0];
          keep = !block_out_of_band(v, c0, c1, c2, c3, ib, blod, bpar, berr, low, high);
        }
        pending &= ~__ballot(mine);
      }
    }
    // Queue the blocks that remain (at most 64 entries and 63 + 64 * kBlockRuns runs pending after this).
    const uint64_t kept = __ballot(keep);
    if (kept) {
      const uint32_t nruns = keep ? bspan.y : 0u;
      const uint32_t runs_incl = wave_scan(nruns);
      const uint32_t e = bq_tail + static_cast<uint32_t>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(kept >> 32),
                                                                                  __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(kept), 0u)));
      if (keep) {
        bq[wib][e & (kBlockQueue - 1)] = make_uint4(w.x | (nruns << 24), bspan.x, w.w, __float_as_uint(scale));
        bq_start[wib][e & (kBlockQueue - 1)] = rq_end + runs_incl - nruns;
      }
      bq_tail += static_cast<uint32_t>(__popcll(kept));
      rq_end += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(runs_incl), 63));
      while (rq_end - rq_pos >= 64) run_stage(64);
    }
  });
  if (rq_end != rq_pos) run_stage(rq_end - rq_pos);
  if (tail != head) {
    wave_sync();
    drain(tail - head);
  }
  occ_flush();
}

// ------------------------------------------------------------------ triangle budget (projected-error histogram)
//
// For a view with a triangle budget, the threshold is chosen before its cut: over the clusters an
// instance pass at the band [T[0], T[B-1]] (times the camera scale) would test, each visible cluster
// adds its triangles to the thresholds b where it is in the plain cut, own <= T[b] < parent, i.e. to
// bins [lower_bound(own), lower_bound(parent)) of the ascending table T. The smallest T[b] whose total
// fits the budget becomes the view's error threshold (the largest when none fits).

// lower_bound over the scaled table: the first b with t[b] * scale >= x (bins when none, and for NaN).
// The table is log-spaced, so the log of x estimates b; stepping against the table itself from there
// gives exactly the binary search's answer (t[b] * scale does not decrease with b).
__device__ __forceinline__ uint32_t bin_of(const float* __restrict__ t, uint32_t bins, float scale, float inv, float x) {
  if (x != x || x > t[bins - 1] * scale) return bins;
  if (!(x > t[0] * scale)) return 0;
  const float u = __logf(x / (t[0] * scale)) * inv;
  int b = u > 0.0f ? static_cast<int>(ceilf(fminf(u, static_cast<float>(bins - 1)))) : 1;
  b = b < 1 ? 1 : b;
  while (b > 0 && t[b - 1] * scale >= x) b--;
  while (b < static_cast<int>(bins) && t[b] * scale < x) b++;
  return static_cast<uint32_t>(b);
}

__global__ __launch_bounds__(256) void k_budget_instances(Params p) {
  __shared__ float4 hzb_rows[4][kHzbMaxLevels];  // per wave: the current view's HZB level rows (occlusion tests)
  const uint32_t lane = lane_id(), wib = threadIdx.x / 64;
  const uint32_t queue_count = *p.cell_queue_count;
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64), wrank = xcd_wave_rank();
  const uint32_t q_end = static_cast<uint32_t>(queue_count * (wrank + 1) / waves);
  uint32_t high_water = 0;  // the most items the partition held after one of this wave's appends
  // The current cell's per-instance values (scale, bounding-sphere centre, LOD-sphere centre), per round of
  // 64 instances, kept while consecutive queue entries (the cell's views) name the same cell.
  uint32_t cur_cell = 0xffffffffu;
  float d_scale[2] = {0.0f, 0.0f}, d_center[2][3] = {{0, 0, 0}, {0, 0, 0}}, d_lq[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
  for (uint32_t q = static_cast<uint32_t>(queue_count * wrank / waves); q < q_end; q++) {
    uint2 entry = p.cell_queue[q];
    const uint32_t vi = uniform(entry.x) & kQueueViewMask;
    const bool cell_in_camera = (uniform(entry.x) & kQInCamera) != 0;  // settled for the whole cell by k_cells
    const uint32_t cell = uniform(entry.y);
    const uint32_t part = work_part(p, vi, wrank, waves);  // as in k_instances
    const size_t part_base = static_cast<size_t>(part) * p.work_capacity;
    const ViewDev& v = p.views[vi];
    if (!v.budget) continue;
    const float4 record = p.c_rec[cell];
    const uint32_t slot = uniform(static_cast<uint32_t>(record.y));
    const uint32_t first_instance = uniform(static_cast<uint32_t>(record.z));
    const uint32_t count = uniform(static_cast<uint32_t>(record.w));
    const float4 levers = p.mesh_table[slot * kMeshStride + 3];
    const uint32_t level_count = min(static_cast<uint32_t>(levers.z), kMaxLevels);
    const uint32_t level_base = static_cast<uint32_t>(levers.w);
    const float4 object_sphere = p.mesh_table[slot * kMeshStride + 0];
    const float4 lod_bounds = p.mesh_table[slot * kMeshStride + 1];
    const float4 flags = p.mesh_table[slot * kMeshStride + 5];
    const bool occlusion_on = flags.z > 0.5f;
    const float min_pixels = fmaxf(v.min_pixel_radius, levers.y);
    if (occlusion_on && v.hzb_valid) {  // wave-uniform: the level table into the wave's LDS rows
      reinterpret_cast<float*>(hzb_rows[wib])[lane] = p.hzbs[v.slot][lane];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (cell != cur_cell) {  // wave-uniform: the cell's instance values, as k_instances derives them
      cur_cell = cell;
#pragma unroll
      for (uint32_t r = 0; r < 2; r++) {
        const bool valid = r * 64 + lane < count;
        const uint32_t instance = first_instance + r * 64 + lane;
        float4 c0 = make_float4(0, 0, 0, 0), c1 = c0, c2 = c0, c3 = make_float4(0, 0, 0, 1);
        if (valid) {
          c0 = p.i_col0[instance];
          c1 = p.i_col1[instance];
          c2 = p.i_col2[instance];
          c3 = p.i_col3[instance];
        }
        d_scale[r] = fmaxf(len3(c0.x, c0.y, c0.z), fmaxf(len3(c1.x, c1.y, c1.z), len3(c2.x, c2.y, c2.z)));
        const float4 center = mat_mul(c0, c1, c2, c3, object_sphere.x, object_sphere.y, object_sphere.z, 1.0f);
        d_center[r][0] = center.x;
        d_center[r][1] = center.y;
        d_center[r][2] = center.z;
        const float4 lq = mat_mul(c0, c1, c2, c3, lod_bounds.x, lod_bounds.y, lod_bounds.z, 1.0f);
        d_lq[r][0] = lq.x;
        d_lq[r][1] = lq.y;
        d_lq[r][2] = lq.z;
        d_lq[r][3] = lq.w;
      }
    }
#pragma unroll
    for (uint32_t r = 0; r < 2; r++) {
      if (r * 64 >= count) break;  // wave-uniform
      const bool valid = r * 64 + lane < count;
      const uint32_t instance = first_instance + r * 64 + lane;
      const float scale = d_scale[r];
      const float4 center = make_float4(d_center[r][0], d_center[r][1], d_center[r][2], 1.0f);
      const float radius = object_sphere.w * scale;
      bool keep = valid && (v.frustum == 0 || cell_in_camera || frustum_distance(v, center.x, center.y, center.z, 0, false) >= -radius);
      const float4 vc = mat_mul16(v.view, center.x, center.y, center.z, 1.0f);
      const float view_distance = len3(vc.x, vc.y, vc.z);
      keep = keep && !(view_distance > radius && (radius * v.proj_scale) / view_distance < min_pixels);
      keep = keep && !((view_distance - radius) > levers.x);
      if (keep && occlusion_on && v.hzb_valid) {
        const float margin = (fabsf(vc.z) * v.budget_high) / v.proj_scale;
        keep = !occluded_by(v, p.hzbs[v.slot], HzbLevelsLds{hzb_rows[wib]}, vc.x, vc.y, vc.z, radius + margin);
      }
      const float4 lv = mat_mul16(v.view, d_lq[r][0], d_lq[r][1], d_lq[r][2], d_lq[r][3]);
      const float d = len3(lv.x, lv.y, lv.z);
      const float rr = lod_bounds.w * scale;
      const float pmin = (scale * v.proj_scale) / fmaxf(d - rr, v.near_plane);
      const float pmax = (scale * v.proj_scale) / fmaxf(d + rr, v.near_plane);
      LevelResult lr{0xffffffffu, 0u, 0u, 0u};
      if (__ballot(keep)) lr = select_band(p, level_base, level_count, v.budget_low, v.budget_high, pmin, pmax);
      const bool work = keep && lr.last > lr.first;
      if (!__ballot(work)) continue;
      // The item covers the LOD blocks of [first, last) (level ranges hold whole runs, runs and blocks never
      // cross a level's bounds).
      const uint32_t run_lo = work ? p.r_block[p.m_run[lr.first]] : 0u, run_hi = work ? p.r_block[p.m_run[lr.last - 1]] + 1 : 0u;
      const uint32_t len = run_hi - run_lo;
      const uint32_t rank = static_cast<uint32_t>(__popcll(__ballot(work) & lanes_below()));
      const uint32_t items = static_cast<uint32_t>(__popcll(__ballot(work)));
      const uint32_t task_incl = wave_scan(len);
      const uint32_t tasks = __shfl(task_incl, 63, 64);
      unsigned long long old = 0;
      if (lane == 0) old = atomicAdd(p.work_counter + part * kWorkCounterSpacing, (static_cast<unsigned long long>(items) << kTaskBits) | tasks);
      const uint32_t old_lo = __shfl(static_cast<uint32_t>(old), 0, 64);
      const uint32_t old_hi = __shfl(static_cast<uint32_t>(old >> 32), 0, 64);
      const uint64_t o = (static_cast<uint64_t>(old_hi) << 32) | old_lo;
      const uint32_t item_base = static_cast<uint32_t>(o >> kTaskBits);
      const uint64_t task_base = o & kTaskMask;
      if (lane == 0 && (task_base + tasks >= kTaskLimit || item_base + items > p.work_capacity)) atomicOr(&p.flags[3], 1u);
      high_water = max(high_water, item_base + items);
      if (work && item_base + rank < p.work_capacity) {
        const uint32_t tag = slot | (occlusion_on ? kTagOcclusion : 0u) | (vi << kTagViewShift);
        const size_t g = part_base + item_base + rank;
        p.work[g] = make_uint4(instance, run_lo, run_hi, tag);
        p.work_start[g] = task_base + task_incl - len;
        p.work_scale[g] = scale;
      }
      wave_count_add(metric_at(p, v.slot, M_BUDGET_TESTS), work ? lr.last - lr.first : 0u, work);
    }
  }
  if (lane == 0 && high_water) atomicMax(&p.flags[4], high_water);
}

// The budget histogram over LOD blocks, in stages per wave as the meshlet pass: (1) a lane per block of the
// budget items; on affine scenes the block test against the band [T[0], T[B-1]] times the camera scale
// drops blocks none of whose runs can add to a bin (every run's own error above the band, so its first
// bin is B, or every parent's at or below it, so its last is 0); the runs of the other blocks are queued
// per wave in LDS. (2) A lane per queued run, 64 at a time: its own and parent errors give the bins [a, e)
// of the thresholds where its meshlets are in the plain cut; the meshlets of runs with a < e go to a
// per-wave LDS ring. (3) A lane per ring meshlet, 64 at a time: each visible one adds its triangles to bins
// [a, e) (its inputs loaded before the per-view loop, the view's HZB level rows copied to LDS).
template <bool kAffine>
__global__ __launch_bounds__(256) void k_budget_meshlets(Params p) {
  // The LDS histograms (dynamic LDS). With p.budget_block_hist: one per launched view, shared by the block's
  // waves, 32-bit two's-complement partial sums added with LDS atomics and moved to the global histograms by
  // exchange (any wave, any time: no add is lost), after every kBudgetFlushMeshlets meshlets a wave adds and
  // at the end. Otherwise, per wave: the histogram of the wave's current view (bins + 1 64-bit words),
  // flushed when the view changes and at the end; lanes of another view add to global memory directly.
  extern __shared__ unsigned long long budget_lds[];
  // Per wave: meshlets of runs that add to some bin, (instance, meshlet, tag, instance scale bits), and their
  // bins a | e << 16.
  __shared__ uint4 ring[4][kQueue];
  __shared__ uint32_t ring_bins[4][kQueue];
  __shared__ float4 hzb_rows[4][kHzbMaxLevels];
  // Per wave: blocks that passed and the position of each one's first run in the stream of queued runs
  // (as in the meshlet pass).
  __shared__ uint4 bq[4][kBlockQueue];
  __shared__ uint32_t bq_start[4][kBlockQueue];
  const uint32_t wib = threadIdx.x / 64, lane = lane_id();
  const uint32_t bins = p.budget_bins;
  const bool block_hist = p.budget_block_hist != 0;  // grid-uniform
  unsigned long long* hist = budget_lds + static_cast<size_t>(wib) * (bins + 1);
  uint32_t* block_words = reinterpret_cast<uint32_t*>(budget_lds);
  const uint32_t block_count = p.view_count * (bins + 1);
  if (block_hist) {
    for (uint32_t i = threadIdx.x; i < block_count; i += blockDim.x) block_words[i] = 0;
    __syncthreads();
  } else {
    for (uint32_t i = lane; i <= bins; i += 64) hist[i] = 0;
  }
  uint32_t added = 0;  // meshlets this wave added to the block histograms since its last flush
  auto block_flush = [&]() __attribute__((always_inline)) {
    for (uint32_t i = lane; i < block_count; i += 64) {
      const int32_t h = static_cast<int32_t>(atomicExch(&block_words[i], 0u));
      if (h) atomicAdd(&p.budget_hist[i], static_cast<unsigned long long>(static_cast<int64_t>(h)));
    }
    added = 0;
  };
  uint32_t current = 0xffffffffu;
  auto flush = [&]() __attribute__((always_inline)) {
    if (current != 0xffffffffu) {
      unsigned long long* g = p.budget_hist + static_cast<size_t>(current) * (bins + 1);
      for (uint32_t i = lane; i <= bins; i += 64) {
        const unsigned long long h = hist[i];
        if (h) atomicAdd(&g[i], h);
        hist[i] = 0;
      }
    }
  };
  uint32_t head = 0, tail = 0;  // ring positions (wave-uniform, running freely)
  uint32_t bq_head = 0, bq_tail = 1, rq_pos = 0, rq_end = 0;
  if (lane == 0) {
    bq[wib][0] = make_uint4(0, 0, 0, 0);
    bq_start[wib][0] = 0;
  }
  auto wave_sync = [&]() __attribute__((always_inline)) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  // Tests the next n (<= 64) ring meshlets, one per lane, and adds the visible ones' triangles to their bins.
  auto drain = [&](uint32_t n) __attribute__((always_inline)) {
    const bool active = lane < n;
    const uint4 c = active ? ring[wib][(head + lane) & (kQueue - 1)] : make_uint4(0, 0, 0, 0);
    const uint32_t ab = active ? ring_bins[wib][(head + lane) & (kQueue - 1)] : 0u;
    const uint32_t a = ab & 0xffffu, e = ab >> 16;
    const uint32_t instance = c.x, meshlet = c.y, tag = c.z, vi = tag >> kTagViewShift;
    const float scale = __uint_as_float(c.w);
    const bool occ_on = (tag & kTagOcclusion) != 0;
    uint32_t meshlet_tris = 0;
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0, c3 = c0, cs = c0;
    if (active) {
      meshlet_tris = ld32(p.m_tris, meshlet);
      c0 = ld32(p.i_col0, instance);
      c1 = ld32(p.i_col1, instance);
      c2 = ld32(p.i_col2, instance);
      c3 = ld32(p.i_col3, instance);
      cs = ld32(p.m_cull, meshlet);
    }
    bool visible = false;
    for (uint64_t pending = __ballot(active); pending;) {  // one launched view at a time (uniform constants)
      const uint32_t vu = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vi), __ffsll(static_cast<unsigned long long>(pending)) - 1));
      const bool mine = active && vi == vu;
      const auto& v = scalar(p.views)[uniform(vu)];
      if (v.hzb_valid && __ballot(mine && occ_on)) {
        reinterpret_cast<float*>(hzb_rows[wib])[lane] = scalar(p.hzbs)[v.slot][lane];
        wave_sync();
      }
      if (mine) {
        const float4 center = mat_mul(c0, c1, c2, c3, cs.x, cs.y, cs.z, 1.0f);
        visible = v.frustum == 0 || frustum_distance(v, center.x, center.y, center.z, 0, false) >= -(cs.w * scale);
        if (visible && occ_on && v.hzb_valid) {
          const float4 vc = mat_mul16(v.view, center.x, center.y, center.z, 1.0f);
          const float margin = (fabsf(vc.z) * v.budget_high) / v.proj_scale;
          visible = !occluded_by(v, scalar(p.hzbs)[v.slot], HzbLevelsLds{hzb_rows[wib]}, vc.x, vc.y, vc.z, cs.w * scale + margin);
        }
      }
      pending &= ~__ballot(mine);
    }
    head += n;
    const uint32_t tris = visible ? meshlet_tris : 0u;
    const bool add = tris != 0;
    if (block_hist) {
      if (add) {
        atomicAdd(&block_words[vi * (bins + 1) + a], tris);
        atomicAdd(&block_words[vi * (bins + 1) + e], 0u - tris);
      }
      added += n;
      if (added >= kBudgetFlushMeshlets) block_flush();
      return;
    }
    // Switch the wave's LDS histogram to the first contributing view when none of the lanes adds to the current one.
    const uint64_t contrib = __ballot(add);
    if (contrib) {
      const uint32_t lead_view = __shfl(vi, __ffsll(static_cast<unsigned long long>(contrib)) - 1, 64);
      if (__ballot(add && vi == current) == 0ull && lead_view != current) {
        flush();
        current = lead_view;
      }
    }
    if (add) {
      if (vi == current) {
        atomicAdd(&hist[a], static_cast<unsigned long long>(tris));
        atomicAdd(&hist[e], static_cast<unsigned long long>(0) - tris);
      } else {
        unsigned long long* g = p.budget_hist + static_cast<size_t>(vi) * (bins + 1);
        atomicAdd(&g[a], static_cast<unsigned long long>(tris));
        atomicAdd(&g[e], static_cast<unsigned long long>(0) - tris);
      }
    }
  };

  // Bins the next n (<= 64) queued runs, one per lane; the meshlets of runs that add to a bin go to the ring
  // (the lanes' k-th meshlets together), which is drained whenever 64 wait (at most 63 + 64 are queued).
  auto run_stage = [&](uint32_t n) __attribute__((always_inline)) {
    wave_sync();
    const uint32_t e_next = bq_head + 1 + lane;
    const bool present = lane < bq_tail - bq_head - 1;
    const uint32_t dist = present ? bq_start[wib][e_next & (kBlockQueue - 1)] - rq_pos : 0xffffffffu;
    const uint64_t bit = dist < 64 ? 1ull << dist : 0ull;
    const uint64_t starts = (static_cast<uint64_t>(wave_or(static_cast<uint32_t>(bit >> 32))) << 32) | wave_or(static_cast<uint32_t>(bit));
    const bool valid = lane < n;
    const uint32_t ent = bq_head + static_cast<uint32_t>(__popcll(starts & ((2ull << lane) - 1ull)));
    uint4 q = make_uint4(0, 0, 0, 0);
    uint32_t first_pos = 0;
    if (valid) {
      q = bq[wib][ent & (kBlockQueue - 1)];
      first_pos = bq_start[wib][ent & (kBlockQueue - 1)];
    }
    const uint32_t instance = q.x & 0xffffffu, tag = q.z, run = q.y + (rq_pos + lane - first_pos);
    const float scale = __uint_as_float(q.w);
    const uint32_t vi = tag >> kTagViewShift;
    float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0, c3 = c0, lod_sphere = c0, parent_sphere = c0;
    float2 err = make_float2(0.0f, 0.0f);
    uint2 span = make_uint2(0, 0);
    if (valid) {
      c0 = ld32(p.i_col0, instance);
      c1 = ld32(p.i_col1, instance);
      c2 = ld32(p.i_col2, instance);
      c3 = ld32(p.i_col3, instance);
      err = ld32(p.r_err, run);
      lod_sphere = ld32(p.r_lod, run);
      parent_sphere = ld32(p.r_parent, run);
      span = ld32(p.r_span, run);
    }
    uint32_t a = 0, e = 0;
    for (uint64_t pending = __ballot(valid); pending;) {  // one launched view at a time (uniform constants)
      const uint32_t vu = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vi), __ffsll(static_cast<unsigned long long>(pending)) - 1));
      const bool mine = valid && vi == vu;
      if (mine) {
        const auto& v = scalar(p.views)[uniform(vu)];
        const float own = projected_error<kAffine>(v, c0, c1, c2, c3, scale, lod_sphere, err.x);
        const float parent = projected_error<kAffine>(v, c0, c1, c2, c3, scale, parent_sphere, err.y);
        const float* t = p.budget_thresholds + static_cast<size_t>(uniform(vu)) * bins;
        a = bin_of(t, bins, v.budget_scale, v.budget_inv, own);
        e = bin_of(t, bins, v.budget_scale, v.budget_inv, parent);
      }
      pending &= ~__ballot(mine);
    }
    // Entries whose runs are all binned now: the new head is the last one starting at or before rq_pos + n.
    const uint64_t upto_n = n >= 64 ? ~0ull : ((2ull << n) - 1ull);
    const bool at_end = __ballot(dist == 64) != 0 && n >= 64;
    bq_head += static_cast<uint32_t>(__popcll(starts & upto_n)) + (at_end ? 1u : 0u);
    rq_pos += n;
    const uint32_t count = valid && a < e ? span.y : 0u;
    for (uint32_t k = 0; k < kMaxRun; k++) {
      const bool has = k < count;
      const uint64_t round = __ballot(has);
      if (!round) break;
      const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(round >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(round), 0u));
      if (has) {
        ring[wib][(tail + below) & (kQueue - 1)] = make_uint4(instance, span.x + k, tag, __float_as_uint(scale));
        ring_bins[wib][(tail + below) & (kQueue - 1)] = a | (e << 16);
      }
      tail += static_cast<uint32_t>(__popcll(round));
      if (tail - head >= 64) {
        wave_sync();
        drain(64);
        wave_sync();
      }
    }
  };

  for_each_task(p, [&](bool valid, uint32_t item, const uint4& w, uint32_t task) __attribute__((always_inline)) {
    if (!__ballot(valid)) return;
    uint2 bspan = make_uint2(0, 0);
    float scale = 0.0f;
    if (valid) {
      bspan = ld32(p.b_span, task);
      scale = p.work_scale[item];
    }
    bool keep = valid;
    if (kAffine) {
      float4 c0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), c1 = c0, c2 = c0, c3 = c0, blod = c0, bpar = c0, berr = c0;
      if (valid) {
        c0 = ld32(p.i_col0, w.x);
        c1 = ld32(p.i_col1, w.x);
        c2 = ld32(p.i_col2, w.x);
        c3 = ld32(p.i_col3, w.x);
        blod = ld32(p.b_lod, task);
        bpar = ld32(p.b_par, task);
        berr = ld32(p.b_err, task);
      }
      const BlockBounds ib = instance_bounds(c0, c1, c2, c3, scale);
      const uint32_t vi = w.w >> kTagViewShift;
      for (uint64_t pending = __ballot(valid); pending;) {
        const uint32_t vu = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vi), __ffsll(static_cast<unsigned long long>(pending)) - 1));
        const bool mine = valid && vi == vu;
        if (mine) {
          const auto& v = scalar(p.views)[uniform(vu)];
          keep = !block_out_of_band(v, c0, c1, c2, c3, ib, blod, bpar, berr, v.budget_low, v.budget_high);
        }
        pending &= ~__ballot(mine);
      }
    }
    // Queue the blocks that remain (at most 64 entries and 63 + 64 * kBlockRuns runs pending after this).
    const uint64_t kept = __ballot(keep);
    if (kept) {
      const uint32_t nruns = keep ? bspan.y : 0u;
      const uint32_t runs_incl = wave_scan(nruns);
      const uint32_t slot = bq_tail + static_cast<uint32_t>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(kept >> 32),
                                                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(kept), 0u)));
      if (keep) {
        bq[wib][slot & (kBlockQueue - 1)] = make_uint4(w.x | (nruns << 24), bspan.x, w.w, __float_as_uint(scale));
        bq_start[wib][slot & (kBlockQueue - 1)] = rq_end + runs_incl - nruns;
      }
      bq_tail += static_cast<uint32_t>(__popcll(kept));
      rq_end += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(runs_incl), 63));
      while (rq_end - rq_pos >= 64) run_stage(64);
    }
  });
  if (rq_end != rq_pos) run_stage(rq_end - rq_pos);
  if (tail != head) {
    wave_sync();
    drain(tail - head);
  }
  if (block_hist) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < block_count; i += blockDim.x) {
      const int32_t h = static_cast<int32_t>(block_words[i]);
      if (h) atomicAdd(&p.budget_hist[i], static_cast<unsigned long long>(static_cast<int64_t>(h)));
    }
  } else {
    flush();
  }
}

// One block per view with a budget: prefix over the bins, pick the threshold, redo the bands.
__global__ __launch_bounds__(64) void k_budget_pick(Params p) {
  ViewDev& v = p.views[blockIdx.x];
  if (blockIdx.x == 0 && threadIdx.x < kWorkParts)  // the regular instance pass refills the work list
    p.work_counter[threadIdx.x * kWorkCounterSpacing] = 0;
  if (!v.budget) return;
  const uint32_t bins = p.budget_bins;
  const unsigned long long* h = p.budget_hist + static_cast<size_t>(blockIdx.x) * (bins + 1);
  const float* t = p.budget_thresholds + static_cast<size_t>(blockIdx.x) * bins;
  const uint32_t lane = lane_id();
  uint64_t carry = 0;
  uint32_t pick = bins - 1;
  bool found = false;
  for (uint32_t base = 0; base < bins && !found; base += 64) {
    const uint32_t b = base + lane;
    const uint64_t d = b < bins ? h[b] : 0ull;
    const uint64_t tri = carry + wave_scan64(d);  // triangles of the plain cut at T[b]
    const uint64_t fits = __ballot(b < bins && tri <= v.budget);
    if (fits) {
      pick = base + static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(fits)) - 1);
      found = true;
    }
    carry = __shfl(tri, 63, 64);
  }
  if (threadIdx.x == 0) {
    v.error_threshold = t[pick];
    v.lod_scale_limit = fmaxf(1.0f, v.max_error_threshold / t[pick]);
    v.budget_bin = pick;
    load_bands(v, state_at(p, v.slot));
    *reinterpret_cast<float*>(metric_at(p, v.slot, M_THRESHOLD)) = t[pick];
    *metric_at(p, v.slot, M_BUDGET_BIN) = pick;
  }
}

// ------------------------------------------------------------------ 4. prefix + capacity (GeometryPool.ts:973-995, 1023-1069)

// `offset`: the first region start (the second phase lays its regions out after the cut's).
__device__ void prefix_cut(const Params& p, uint32_t slot, uint32_t cut, uint32_t offset = 0) {
  const uint32_t lane = lane_id();
  const uint32_t capacity = p.cut_capacity[cut];
  uint32_t* args = p.args + static_cast<size_t>(slot) * p.args_stride;
  uint32_t carry = offset;
  for (uint32_t base = 0; base < p.mesh_slots; base += 64) {
    const uint32_t mesh = base + lane;
    const bool valid = mesh < p.mesh_slots;
    uint32_t c0 = valid ? *mesh_count_at(p, slot, cut, mesh, 0) : 0u;
    uint32_t c1 = valid ? *mesh_count_at(p, slot, cut, mesh, 1) : 0u;
    const uint32_t sum = c0 + c1;
    const uint32_t incl = wave_scan(sum);
    const uint32_t start0 = carry + incl - sum;
    const uint32_t start1 = start0 + c0;
    if (valid) {
      const uint32_t starts[2] = {start0, start1};
      const uint32_t req[2] = {c0, c1};
      for (uint32_t part = 0; part < 2; part++) {
        *(mesh_count_at(p, slot, cut, mesh, part) + 1) = starts[part];
        const uint32_t room = starts[part] < capacity ? capacity - starts[part] : 0u;
        uint32_t* a = args + cut * p.slots * 10 + part * 5 + mesh * 10;
        a[0] = req[part] < room ? req[part] : room;
        a[1] = 1;
        a[2] = starts[part];
        a[3] = 0;
        a[4] = 0;
      }
    }
    carry += __shfl(incl, 63, 64);
  }
  if (lane == 0) {
    *counter_at(p, slot, cut, kRequested) = carry;
    const uint32_t selected = *counter_at(p, slot, cut, kSelected);
    *counter_at(p, slot, cut, kDrawn) = selected < p.max_meshlets ? selected : p.max_meshlets;
  }
}

__device__ float fill_of(const Params& p, uint32_t slot, uint32_t cut) {
  const float requested = static_cast<float>(*counter_at(p, slot, cut, kRequested));
  const float selected = static_cast<float>(*counter_at(p, slot, cut, kSelected));
  return fmaxf(requested / static_cast<float>(p.cut_capacity[cut]), selected / static_cast<float>(p.max_meshlets));
}

// mode 0: prefix then checkNode of retry round `round` (1-based); mode 1: prefix then steerNode.
__global__ __launch_bounds__(128) void k_prefix(Params p, uint32_t mode, uint32_t round) {
  ViewDev& v = p.views[blockIdx.x];
  const uint32_t slot = v.slot;
  const uint32_t cut = threadIdx.x / 64;
  if (mode == 2) {  // second phase: regions after the cut's requested indices; no check or steer
    prefix_cut(p, slot, cut, cut == 0 ? p.counters1[static_cast<size_t>(slot) * p.counter_stride + 2 + kRequested] : 0u);
    return;
  }
  prefix_cut(p, slot, cut);
  __syncthreads();
  if (mode == 0 && blockIdx.x == 0 && threadIdx.x < kWorkParts)  // the retry round's instance pass refills the work list
    p.work_counter[threadIdx.x * kWorkCounterSpacing] = 0;
  if (threadIdx.x != 0) return;
  uint32_t* state = state_at(p, slot);
  if (mode == 0) {
    bool any = false;
    for (uint32_t cut = 0; cut < 2; cut++) {
      const float fill = fill_of(p, slot, cut);
      const float scale = get_state(state, kStateScale + cut);
      set_state(state, kStateRetry + cut, 0.0f);
      v.redo[cut] = 0;
      if (fill > 1.0f && scale < v.lod_scale_limit) {
        set_state(state, kStateScale + cut, fminf(scale * clampf(sqrtf(fill / 0.82f), 1.1f, 4.0f), v.lod_scale_limit));
        set_state(state, kStateRetry + cut, 1.0f);
        v.redo[cut] = 1;
        any = true;
        for (uint32_t k = 0; k < kCutCounters; k++) *counter_at(p, slot, cut, k) = 0;
        if (cut == 0) view_counters(p, slot)[1] = 0;
        if (cut == 0 && v.two_phase) phase_at(p, slot)[P_REJECTS] = phase_at(p, slot)[P_OVERFLOW] = 0;  // recorded again
        for (uint32_t i = 0; i < p.mesh_slots; i++) {
          *mesh_count_at(p, slot, cut, i, 0) = 0;
          *mesh_count_at(p, slot, cut, i, 1) = 0;
        }
      }
    }
    load_bands(v, state);
    if (any) atomicOr(&p.flags[round - 1], 1u);
  } else {
    if (v.two_phase) {  // the band the final round tested with, for the second phase
      uint32_t* pw = phase_at(p, slot);
      pw[P_LOW] = __float_as_uint(v.low[0]);
      pw[P_HIGH] = __float_as_uint(v.high[0]);
      pw[P_FADE] = __float_as_uint(v.fade_scale);
    }
    for (uint32_t cut = 0; cut < 2; cut++) {
      const float fill = fill_of(p, slot, cut);
      float scale = get_state(state, kStateScale + cut);
      if (fill > 0.9f) scale = scale * fminf(sqrtf(fill / 0.82f), 1.03f);
      else if (fill < 0.75f) scale = scale / 1.01f;
      set_state(state, kStateScale + cut, fmaxf(fminf(scale, v.lod_scale_limit), 1.0f));
      set_state(state, kStateFill + cut, fill);
    }
  }
}

// ------------------------------------------------------------------ second phase of two-phase occlusion
//
// The cut of a two-phase view records its camera cut's occlusion rejects (final round): instances
// rejected by the instance test in a bitmap, meshlets rejected in the meshlet pass as (instance, meshlet,
// blend range, mesh slot, scale). Once the caller has drawn the cut and set the slot's new HZB, the
// second phase runs the cells, instance and meshlet passes again on the rejected instances only (camera
// cut, the final round's band, occlusion against the new HZB) and re-tests the rejected meshlets' culling
// spheres against the new HZB (their frustum and LOD tests passed). What is visible now is appended to the
// camera draw list after the cut's entries and laid out in regions after the cut's indices.

// expand_plan[view] = first 64-record chunk of the view's meshlet rejects; expand_plan[views] = the total.
__global__ __launch_bounds__(64) void k_retest_plan(Params p) {
  const uint32_t lane = lane_id();
  unsigned long long carry = 0;
  for (uint32_t base = 0; base < p.view_count; base += 64) {
    const uint32_t vi = base + lane;
    uint32_t chunks = 0;
    if (vi < p.view_count) chunks = (min(phase_at(p, p.views[vi].slot)[P_REJECTS], p.reject_capacity) + 63u) / 64u;
    const uint32_t incl = wave_scan(chunks);
    if (vi < p.view_count) p.expand_plan[vi] = carry + (incl - chunks);
    carry += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
  }
  if (lane == 0) p.expand_plan[p.view_count] = carry;
}

__global__ __launch_bounds__(256) void k_retest(Params p) {
  const uint32_t lane = lane_id();
  const unsigned long long* __restrict__ plan = p.expand_plan;
  const uint64_t total = plan[p.view_count];
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
  uint32_t vi = 0;
  for (uint64_t c = xcd_wave_rank(); c < total; c += waves) {
    while (plan[vi + 1] <= c) vi++;
    const auto& v = scalar(p.views)[uniform(vi)];
    const uint32_t slot = uniform(v.slot);
    const uint32_t n = min(phase_at(p, slot)[P_REJECTS], p.reject_capacity);
    const uint32_t e = static_cast<uint32_t>(c - plan[vi]) * 64u + lane;
    const bool active = e < n;
    const uint4 r = active ? p.reject_meshlets[static_cast<size_t>(slot) * p.reject_capacity + e] : make_uint4(0, 0, 0, 0);
    const uint32_t instance = r.x, meshlet = r.y, fades = r.z & 0xffffu, mslot = r.z >> 16;
    bool visible = active;
    const bool occ_test = active && v.hzb_valid;  // recorded meshlets are of meshes with occlusion on
    if (occ_test) {
      const float4 c0 = p.i_col0[instance], c1 = p.i_col1[instance], c2 = p.i_col2[instance], c3 = p.i_col3[instance];
      const float4 cs = p.m_cull[meshlet];
      const float4 center = mat_mul(c0, c1, c2, c3, cs.x, cs.y, cs.z, 1.0f);
      visible = !occluded_sphere(p, v, true, center.x, center.y, center.z, cs.w * __uint_as_float(r.w));
    }
    const uint32_t tests = static_cast<uint32_t>(__popcll(__ballot(occ_test)));
    if (lane == 0 && tests) atomicAdd(metric_at(p, slot, M_OCCLUSION), static_cast<unsigned long long>(tests));
    const uint32_t tris = visible ? p.m_tris[meshlet] : 0u;
    const uint32_t packed_tris = tris | (fades << 8);
    const uint32_t at = wave_atomic_add_uniform(counter_at(p, slot, 0, kSelected), 1u, visible);
    const bool write = visible && at < p.max_meshlets;
    const uint32_t part = (packed_tris & 0xffff00u) == kFullFadeBits ? 0u : 1u;
    const uint32_t first_index = wave_atomic_add(mesh_count_at(p, slot, 0, mslot, part), tris * 3, write);
    if (write) p.draws[slot][at] = make_uint4(instance, meshlet, first_index, packed_tris);
  }
}

// ------------------------------------------------------------------ 5. expand (expandPass, GeometryPool.ts:997-1021)
//
// Each drawn entry of a cut writes 3 indices per triangle, (draw-list position << 7) | local vertex id,
// at regionStart(cut, mesh slot, part) + firstIndex + 3i, skipping a triangle whose last index does not
// fit the cut's index capacity. The reference runs a workgroup per drawn meshlet. Here the draw lists of
// all launched views' cuts are numbered as one sequence of 64-entry chunks dealt round-robin to
// persistent waves (the draw lists of views clamped by their capacity have little to write, so
// contiguous shares would leave some waves most of the work). A first pass resolves each entry's
// destination; the writing pass then takes a chunk's resolved records with one coalesced load and
// writes the entries that have anything to write, a lane per triangle, several entries at a time.

// The first draw-list entry the expand pass writes: the second phase writes only the entries it appended.
__device__ __forceinline__ uint32_t expand_first(const Params& p, uint32_t slot, uint32_t cut) {
  return p.phase == 2 && cut == 0 ? p.counters1[static_cast<size_t>(slot) * p.counter_stride + 2 + kDrawn] : 0u;
}

// expand_plan[2 * view + cut] = first chunk of that cut's draw list; expand_plan[2 * views] = the total.
__global__ __launch_bounds__(64) void k_expand_plan(Params p) {
  const uint32_t lane = lane_id();
  const uint32_t segments = 2 * p.view_count;
  unsigned long long carry = 0;
  for (uint32_t base = 0; base < segments; base += 64) {
    const uint32_t seg = base + lane;
    uint32_t chunks = 0;
    if (seg < segments) {
      const uint32_t slot = p.views[seg >> 1].slot;
      chunks = (*counter_at(p, slot, seg & 1u, kDrawn) - expand_first(p, slot, seg & 1u) + 63u) / 64u;
    }
    const uint32_t incl = wave_scan(chunks);  // at most 64 x 2^18 (drawn <= 2^24 per cut)
    if (seg < segments) p.expand_plan[seg] = carry + (incl - chunks);
    carry += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
  }
  if (lane == 0) p.expand_plan[segments] = carry;
}

// The segment (view, cut) of chunk c, from a segment at or before it (chunks only grow along a wave).
__device__ __forceinline__ uint32_t segment_of(const unsigned long long* __restrict__ plan, uint32_t seg, uint64_t c) {
  while (plan[seg + 1] <= c) seg++;  // past empty segments
  return seg;
}

// Resolves every drawn entry, a lane per entry, into (first index, first triangle word, triangles to
// write, draw-list position): the gathers of its mesh slot and region start are independent across
// entries here, so they overlap instead of forming a chain per chunk in the writing kernel. The count
// is 0 for an entry wholly past the capacity.
__global__ __launch_bounds__(256) void k_expand_resolve(Params p) {
  const uint32_t lane = lane_id();
  const uint32_t segments = 2 * p.view_count;
  const unsigned long long* __restrict__ plan = p.expand_plan;
  const uint64_t total = plan[segments];
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
  uint32_t seg = 0;
  for (uint64_t c = xcd_wave_rank(); c < total; c += waves) {
    seg = segment_of(plan, seg, c);
    const uint32_t cut = seg & 1u;
    const uint32_t slot = uniform(p.views[seg >> 1].slot);
    const uint32_t drawn = *counter_at(p, slot, cut, kDrawn);
    const uint32_t first = expand_first(p, slot, cut) + static_cast<uint32_t>(c - plan[seg]) * 64u;
    const uint32_t position = first + lane + cut * p.max_meshlets;
    const uint32_t limit = p.cut_capacity[cut] - 2u;  // index + 2 < capacity  <=>  index < limit
    uint4 r = make_uint4(0, 0, 0, position);
    if (first + lane < drawn) {
      const uint4 entry = p.draws[slot][position];
      const uint32_t part = (entry.w & 0xffff00u) == kFullFadeBits ? 0u : 1u;
      const uint32_t region = *(mesh_count_at(p, slot, cut, p.i_slot[entry.x], part) + 1);
      const uint32_t dst = region + entry.z;  // u32 sums, as in the reference
      const uint32_t tris = entry.w & 255u;
      const uint32_t last = dst + 3u * (tris - 1u);
      // Something to write: the first triangle fits (dst < limit), or a later one's index wraps.
      const bool any = tris > 0 && (dst < limit || last < dst);
      r = make_uint4(dst, p.m_first[entry.y] / 3u, any ? tris : 0u, position);
    }
    p.expand_records[c * 64 + lane] = r;
  }
}

__global__ __launch_bounds__(256) void k_expand(Params p) {
  const uint32_t lane = lane_id();
  const uint32_t segments = 2 * p.view_count;
  const unsigned long long* __restrict__ plan = p.expand_plan;
  const uint64_t total = plan[segments];
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x / 64);
  const uint32_t* __restrict__ tri = p.m_tri;
  // Global (not generic) stores; a triangle's 3 indices are one 12-byte store (4-byte aligned).
  struct Triple {
    uint32_t a, b, c;
  };
  uint32_t seg = 0;
  uint64_t c = xcd_wave_rank();
  uint4 next = c < total ? p.expand_records[c * 64 + lane] : make_uint4(0, 0, 0, 0);
  for (; c < total; c += waves) {
    seg = segment_of(plan, seg, c);
    const uint4 r = next;  // this chunk's records; the next chunk's load overlaps its writes
    if (c + waves < total) next = p.expand_records[(c + waves) * 64 + lane];
    uint64_t todo = __ballot(r.z != 0);
    if (!todo) continue;
    const uint32_t cut = seg & 1u;
    const uint32_t slot = uniform(p.views[seg >> 1].slot);
    const uint32_t limit = p.cut_capacity[cut] - 2u;
    VG_GLOBAL uint32_t* out = (VG_GLOBAL uint32_t*)(p.indices[2 * slot + cut]);
    // Up to kBatch entries at a time: their triangle words are loaded together, then all their indices
    // stored. A store counts in the same counter as a load here, so the wait before the stores also
    // waits for the previous batch's stores; batching amortises that round trip over kBatch entries.
    constexpr int kBatch = 16;
    while (todo) {
      uint32_t kk[kBatch], w0[kBatch], w1[kBatch];
#pragma unroll
      for (int g = 0; g < kBatch; g++) {
        kk[g] = todo ? static_cast<uint32_t>(__builtin_ctzll(todo)) : 64u;
        todo &= todo - 1;
        const uint32_t k = kk[g] & 63u;
        const uint32_t f = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.y), static_cast<int>(k)));
        const uint32_t t = kk[g] < 64u ? static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.z), static_cast<int>(k))) : 0u;
        w0[g] = lane < t ? tri[f + lane] : 0u;
        w1[g] = lane + 64u < t ? tri[f + lane + 64u] : 0u;
      }
#pragma unroll
      for (int g = 0; g < kBatch; g++) {
        if (kk[g] >= 64u) break;
        const uint32_t k = kk[g];
        const uint32_t d = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.x), static_cast<int>(k)));
        const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.z), static_cast<int>(k)));
        const uint32_t base = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.w), static_cast<int>(k))) << kLocalVertexBits;
        auto put = [&](uint32_t i, uint32_t w) {
          const uint32_t index = d + 3u * i;  // the triangle's first index
          if (index < limit) {
            VG_GLOBAL Triple* q = (VG_GLOBAL Triple*)(out + index);
            q->a = base | (w & 255u);
            q->b = base | ((w >> 8) & 255u);
            q->c = base | ((w >> 16) & 255u);
          }
        };
        if (lane < t) put(lane, w0[g]);
        if (lane + 64u < t) put(lane + 64u, w1[g]);
        if (t > 128u) {  // meshlets of more than 128 triangles (up to 255)
          const uint32_t f = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(r.y), static_cast<int>(k)));
          for (uint32_t i = lane + 128u; i < t; i += 64u) put(i, tri[f + i]);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ HZB (OcclusionCulling.ts:302-350)

// Reduces levels base+1 .. base+5 (those that exist) over 32 x 32 texel tiles of level `base`.
// One block per 32 x 32 tile of level `base` of `hzb`; a block outside the level (a grid sized for the largest
// of several HZBs) returns at once.
__device__ __forceinline__ void hzb_pass(float* hzb, uint32_t base) {
  __shared__ float tile[2][32 * 32];
  const uint32_t levels = static_cast<uint32_t>(hzb[kHzbLevelCountSlot]);
  const uint32_t W = static_cast<uint32_t>(hzb[base * 4 + 1]), H = static_cast<uint32_t>(hzb[base * 4 + 2]);
  const uint32_t src_offset = static_cast<uint32_t>(hzb[base * 4]);
  const uint32_t tx = blockIdx.x * 32, ty = blockIdx.y * 32;
  if (base >= levels || tx >= W || ty >= H) return;  // block-uniform (header rows past the last level are 0)
  for (uint32_t i = threadIdx.x; i < 32 * 32; i += blockDim.x) {
    const uint32_t x = min(tx + (i & 31), W - 1), y = min(ty + (i >> 5), H - 1);
    tile[0][i] = hzb[src_offset + y * W + x];
  }
  __syncthreads();
  uint32_t cur = 0;
  for (uint32_t l = base + 1, side = 16; l < levels && l <= base + 5; l++, side >>= 1) {
    const uint32_t sw = static_cast<uint32_t>(hzb[(l - 1) * 4 + 1]), sh = static_cast<uint32_t>(hzb[(l - 1) * 4 + 2]);
    const uint32_t lw = static_cast<uint32_t>(hzb[l * 4 + 1]), lh = static_cast<uint32_t>(hzb[l * 4 + 2]);
    const uint32_t off = static_cast<uint32_t>(hzb[l * 4]);
    const uint32_t ox = tx >> (l - base), oy = ty >> (l - base);       // tile origin at level l
    const uint32_t px = tx >> (l - 1 - base), py = ty >> (l - 1 - base);  // tile origin at level l - 1
    for (uint32_t i = threadIdx.x; i < side * side; i += blockDim.x) {
      const uint32_t lx = i % side, ly = i / side;
      const uint32_t x = ox + lx, y = oy + ly;
      float value = 0.0f;
      if (x < lw && y < lh) {
        const uint32_t x0 = x * 2, y0 = y * 2;
        const uint32_t x1 = min(x0 + 1, sw - 1), y1 = min(y0 + 1, sh - 1);
        const float* s = tile[cur];
        const float a = fmaxf(s[(y0 - py) * 32 + (x0 - px)], s[(y0 - py) * 32 + (x1 - px)]);
        const float b = fmaxf(s[(y1 - py) * 32 + (x0 - px)], s[(y1 - py) * 32 + (x1 - px)]);
        value = fmaxf(a, b);
        hzb[off + y * lw + x] = value;
      }
      tile[cur ^ 1][ly * 32 + lx] = value;
    }
    __syncthreads();
    cur ^= 1;
  }
}

__global__ __launch_bounds__(256) void k_hzb_pass(float* hzb, uint32_t base) { hzb_pass(hzb, base); }

// words[0, n) <- 0 (an HZB's texel area before vg_hzb_set writes its image).
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* words, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0u;
}

// Population count of a bitmap into *out (measurement only).
__global__ __launch_bounds__(256) void k_popcount(const uint32_t* bits, uint32_t words, unsigned long long* out) {
  uint32_t n = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) n += __popc(bits[i]);
  wave_count_add(out, n, n != 0);
}

// ------------------------------------------------------------------ arithmetic probe (tests)

// Evaluates the operations the cut's decisions depend on, for comparison with the host.
__global__ void k_arith_probe(const float* a, const float* b, float* out, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a[i], y = b[i];
  out[i * 5 + 0] = x / y;
  out[i * 5 + 1] = __builtin_sqrtf(fabsf(x));
  out[i * 5 + 2] = vg_logf(fabsf(x));
  out[i * 5 + 3] = rintf(x);
  out[i * 5 + 4] = static_cast<float>(__float_as_uint(y) >> 3);
}

// Wave-scan probe: for each group of 64 values, the inclusive prefix sum and the group's sum.
__global__ __launch_bounds__(64) void k_scan_probe(const uint32_t* in, uint32_t* out, uint32_t groups) {
  const uint32_t g = blockIdx.x;
  if (g >= groups) return;
  const uint32_t x = in[g * 64 + threadIdx.x];
  out[g * 128 + threadIdx.x] = wave_scan(x);
  out[g * 128 + 64 + threadIdx.x] = wave_sum(x);
}

}  // namespace vg
