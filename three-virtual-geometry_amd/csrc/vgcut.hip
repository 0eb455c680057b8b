// Host runtime of the cut behind include/vgcut.h: contexts, resident scenes (GeometryPool.ts:213-512),
// launch sequences and HIP-graph replay (VirtualGeometry.ts:392-507), HZB builds
// (OcclusionCulling.ts:244-352). One translation unit with the kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/vgcut.h"
#include "vgcut_kernels.hip"
#include "vgraster_kernels.hip"
#include "vgcut_blocks.h"
#include "vgcut_cells.h"
#include "vgcut_host.h"

using namespace vg;
using namespace vgi;

extern "C" int vg_budget_thresholds(float lo, float hi, uint32_t bins, float* out);

// Scratch of vg_cut_static: the launch buffers a cut needs beyond its inputs and outputs (view records,
// the (view, cell) queue, the work list with its counters and flags, budget tables), in one array the
// code object holds on every device it is loaded on, for the life of the process. Host code never uses
// its name as an address: vg_create looks up its device address (hipGetSymbolAddress) with the context's
// device current and keeps it in the context.
constexpr size_t kStaticScratchBytes = 64ull << 20;
constexpr uint32_t kStaticViews = 64;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace vgi {

int fail(vg_ctx* ctx, int code, const std::string& message) {
  if (ctx) ctx->error = message;
  return code;
}

int hip_fail(vg_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, VG_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define VG_HIP(ctx, expr)                                 \
  do {                                                    \
    hipError_t e_ = (expr);                               \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

// Grows `b` to at least `bytes` (contents are not kept). Returns 0 or an error code.
int ensure(vg_scene* s, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return 0;
  if (b.ptr) {
    (void)hipFree(b.ptr);
    s->device_bytes -= b.bytes;
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t n = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, n);
  if (e != hipSuccess) return fail(s->ctx, VG_E_NOMEM, std::string("hipMalloc of ") + std::to_string(n) + " bytes failed");
  b.bytes = n;
  s->device_bytes += n;
  s->layout_version++;
  return 0;
}

// Lane scratch: the same, accounted to the lane.
int ensure_lane(vg_ctx* ctx, Lane* lane, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return 0;
  if (b.ptr) {
    (void)hipFree(b.ptr);
    lane->device_bytes -= b.bytes;
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t n = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.ptr, n);
  if (e != hipSuccess) return fail(ctx, VG_E_NOMEM, std::string("hipMalloc of ") + std::to_string(n) + " bytes failed");
  b.bytes = n;
  lane->device_bytes += n;
  lane->layout_version++;
  return 0;
}

void release_lane(Lane* lane, DevBuf& b) {
  if (b.ptr) {
    (void)hipFree(b.ptr);
    lane->device_bytes -= b.bytes;
  }
  b.ptr = nullptr;
  b.bytes = 0;
}

int lane_init(vg_ctx* ctx, Lane* lane, hipStream_t stream) {
  if (stream) {
    lane->stream = stream;
    lane->owns_stream = false;
  } else {
    VG_HIP(ctx, hipStreamCreateWithFlags(&lane->stream, hipStreamNonBlocking));
    lane->owns_stream = true;
  }
  int rc = 0;
  if ((rc = ensure_lane(ctx, lane, lane->d_cell_queue_count, 64)) ||
      (rc = ensure_lane(ctx, lane, lane->d_work_counter, sizeof(unsigned long long) * kWorkParts * kWorkCounterSpacing)) ||
      (rc = ensure_lane(ctx, lane, lane->d_flags, 64)))
    return rc;
  VG_HIP(ctx, hipMemset(lane->d_flags.ptr, 0, lane->d_flags.bytes));
  for (int r = 0; r < Lane::kRing; r++) VG_HIP(ctx, hipEventCreateWithFlags(&lane->ring_copied[r], hipEventDisableTiming));
  return 0;
}

void lane_destroy(Lane* lane) {
  if (lane->stream) (void)hipStreamSynchronize(lane->stream);
  if (lane->graph_exec) (void)hipGraphExecDestroy(lane->graph_exec);
  if (lane->graph) (void)hipGraphDestroy(lane->graph);
  lane->graph_exec = nullptr;
  lane->graph = nullptr;
  DevBuf* bufs[] = {&lane->d_views, &lane->d_cell_queue_count, &lane->d_work_counter, &lane->d_flags, &lane->d_budget_thresholds,
                    &lane->d_budget_hist, &lane->d_expand_plan, &lane->d_expand_records};
  for (DevBuf* b : bufs) release_lane(lane, *b);
  if (!lane->borrowed) {
    DevBuf* t[] = {&lane->d_cell_queue, &lane->d_work, &lane->d_work_start, &lane->d_work_scale};
    for (DevBuf* b : t) release_lane(lane, *b);
  }
  for (int r = 0; r < Lane::kRing; r++) {
    if (lane->h_ring[r]) (void)hipHostFree(lane->h_ring[r]);
    if (lane->h_ring_thresholds[r]) (void)hipHostFree(lane->h_ring_thresholds[r]);
    if (lane->ring_copied[r]) (void)hipEventDestroy(lane->ring_copied[r]);
    lane->h_ring[r] = nullptr;
    lane->h_ring_thresholds[r] = nullptr;
    lane->ring_copied[r] = nullptr;
  }
  if (lane->owns_stream && lane->stream) (void)hipStreamDestroy(lane->stream);
  lane->stream = nullptr;
}

void transient_sizes(const vg_scene* s, uint32_t count, size_t* queue_bytes, size_t* work_items) {
  // Every (view, cell) can be queued once. A partition of the work list takes the items of at most
  // ceil(entries / kWorkParts) + 2 queue entries (partitions by wave rank, launches of fewer than kViewPartsMin
  // views) or of the entries of ceil(count / kWorkParts) views (partitions by view), each entry at most
  // 2 cuts x 128 instances, nor more than all of those views' (instance, cut) pairs.
  const size_t cells = std::max<size_t>(s->c_sphere.size(), 1);
  *queue_bytes = sizeof(uint2) * cells * count;
  const uint64_t entries = static_cast<uint64_t>(cells) * count;
  const bool by_view = count >= kViewPartsMin;
  const uint64_t views_per_part = by_view ? (count + kWorkParts - 1) / kWorkParts : count;
  const uint64_t entries_per_part = by_view ? static_cast<uint64_t>(cells) * views_per_part : (entries + kWorkParts - 1) / kWorkParts + 2;
  const uint64_t per_part = std::min<uint64_t>(entries_per_part * 2 * kCellSize,
                                               static_cast<uint64_t>(s->i_col[0].size()) * 2 * views_per_part + 64);
  *work_items = std::min<uint64_t>(per_part, kMaxWorkItems / kWorkParts) * kWorkParts;
}

void release(vg_scene* s, DevBuf& b) {
  if (b.ptr) {
    (void)hipFree(b.ptr);
    s->device_bytes -= b.bytes;
  }
  b.ptr = nullptr;
  b.bytes = 0;
}

template <typename T>
int upload(vg_scene* s, DevBuf& b, const std::vector<T>& v) {
  if (int rc = ensure(s, b, sizeof(T) * std::max<size_t>(v.size(), 1))) return rc;
  if (!v.empty()) VG_HIP(s->ctx, hipMemcpyAsync(b.ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s->ctx->stream));
  return 0;
}

uint32_t spread_bits10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// mortonOrder + radixOrder, GeometryPool.ts:1263-1311: caller index of each GPU slot.
std::vector<uint32_t> morton_order(const float* matrices, uint32_t n, const float* sphere) {
  std::vector<double> centers(static_cast<size_t>(n) * 3);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double s[4];
  for (uint32_t i = 0; i < n; i++) {
    instance_sphere(matrices + static_cast<size_t>(i) * kInstanceStride, sphere, s);
    for (int a = 0; a < 3; a++) {
      centers[i * 3 + a] = s[a];
      lo[a] = std::min(lo[a], s[a]);
      hi[a] = std::max(hi[a], s[a]);
    }
  }
  const double extent = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), std::max(hi[2] - lo[2], 1e-9));
  std::vector<uint32_t> codes(n);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
      double v = std::floor(((centers[i * 3 + a] - lo[a]) / extent) * 1024);
      q[a] = static_cast<uint32_t>(std::min(1023.0, std::max(0.0, v)));
    }
    codes[i] = spread_bits10(q[0]) | (spread_bits10(q[1]) << 1) | (spread_bits10(q[2]) << 2);
  }
  std::vector<uint32_t> order(n), next(n), counts(1u << 15);
  for (uint32_t i = 0; i < n; i++) order[i] = i;
  for (uint32_t shift : {0u, 15u}) {
    std::fill(counts.begin(), counts.end(), 0u);
    for (uint32_t i = 0; i < n; i++) counts[(codes[i] >> shift) & 0x7fffu]++;
    uint32_t sum = 0;
    for (auto& c : counts) {
      uint32_t v = c;
      c = sum;
      sum += v;
    }
    for (uint32_t i = 0; i < n; i++) {
      uint32_t index = order[i];
      next[counts[(codes[index] >> shift) & 0x7fffu]++] = index;
    }
    order.swap(next);
  }
  return order;
}

void write_mesh_record(vg_scene* s, uint32_t slot) {
  const MeshHost& m = s->meshes[slot];
  float4* r = &s->mesh_table[slot * kMeshStride];
  r[0] = make_float4(m.bounding[0], m.bounding[1], m.bounding[2], m.bounding[3]);
  r[1] = make_float4(m.lod_bounds[0], m.lod_bounds[1], m.lod_bounds[2], m.lod_bounds[3]);
  r[2] = make_float4(m.cull_bounds[0], m.cull_bounds[1], m.cull_bounds[2], m.cull_bounds[3]);
  r[4] = make_float4(1, 1, 0, 0);
}

int upload_scene(vg_scene* s) {
  int rc = 0;
  if (s->dirty_meshlets) {
    if ((rc = upload(s, s->d_cull, s->m_cull)) || (rc = upload(s, s->d_lod, s->m_lod)) || (rc = upload(s, s->d_parent, s->m_parent)) ||
        (rc = upload(s, s->d_err, s->m_err)) || (rc = upload(s, s->d_tris, s->m_tris)) || (rc = upload(s, s->d_run, s->m_run)) ||
        (rc = upload(s, s->d_rspan, s->r_span)) || (rc = upload(s, s->d_rlod, s->r_lod)) ||
        (rc = upload(s, s->d_rparent, s->r_parent)) || (rc = upload(s, s->d_rerr, s->r_err)) ||
        (rc = upload(s, s->d_rblock, s->r_block)) || (rc = upload(s, s->d_bspan, s->b_span)) ||
        (rc = upload(s, s->d_blod, s->b_lod)) || (rc = upload(s, s->d_bpar, s->b_par)) || (rc = upload(s, s->d_berr, s->b_err)) ||
        (rc = upload(s, s->d_level_blocks, s->level_blocks)) ||
        (rc = upload(s, s->d_first, s->m_first)) || (rc = upload(s, s->d_tri, s->m_tri)) ||
        (rc = upload(s, s->d_vlist, s->m_vlist)))
      return rc;
    {
      std::vector<float4> vpos(s->m_mverts.size());  // positions in meshlet-vertex order (ids checked by vg_mesh_add)
      for (size_t i = 0; i < vpos.size(); i++) vpos[i] = s->m_verts[s->m_mverts[i]];
      if ((rc = upload(s, s->d_vpos, vpos))) return rc;
      VG_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));  // vpos is a temporary
    }
    auto finite4 = [](const float4& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z) && std::isfinite(v.w); };
    s->finite_runs = std::all_of(s->r_lod.begin(), s->r_lod.end(), finite4) && std::all_of(s->r_parent.begin(), s->r_parent.end(), finite4);
    s->dirty_meshlets = false;
  }
  if (s->dirty_table) {
    if ((rc = upload(s, s->d_mesh_table, s->mesh_table))) return rc;
    s->dirty_table = false;
  }
  if (s->dirty_instances) {
    for (int k = 0; k < 4; k++)
      if ((rc = upload(s, s->d_col[k], s->i_col[k]))) return rc;
    std::vector<uint32_t> islot(s->i_slot.size());
    for (size_t i = 0; i < islot.size(); i++) islot[i] = static_cast<uint32_t>(s->i_slot[i]);
    if ((rc = upload(s, s->d_islot, islot))) return rc;
    VG_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));  // islot is a temporary
    bool affine = true;
    for (size_t i = 0; affine && i < s->i_col[0].size(); i++) {
      const float4 c[4] = {s->i_col[0][i], s->i_col[1][i], s->i_col[2][i], s->i_col[3][i]};
      affine = c[0].w == 0.0f && c[1].w == 0.0f && c[2].w == 0.0f && c[3].w == 1.0f;
      for (int k = 0; affine && k < 4; k++) affine = std::isfinite(c[k].x) && std::isfinite(c[k].y) && std::isfinite(c[k].z);
    }
    s->affine_instances = affine;
    s->dirty_instances = false;
  }
  if (s->dirty_cells) {
    s->c_sphere.clear();
    s->c_rec.clear();
    s->c_bound.clear();
    for (const MeshHost& m : s->meshes) {
      if (!m.used) continue;
      s->c_sphere.insert(s->c_sphere.end(), m.cell_sphere.begin(), m.cell_sphere.end());
      s->c_rec.insert(s->c_rec.end(), m.cell_rec.begin(), m.cell_rec.end());
      s->c_bound.insert(s->c_bound.end(), m.cell_bound.begin(), m.cell_bound.end());
    }
    if ((rc = upload(s, s->d_csphere, s->c_sphere)) || (rc = upload(s, s->d_crec, s->c_rec)) || (rc = upload(s, s->d_cbound, s->c_bound)))
      return rc;
    s->dirty_cells = false;
  }
  return 0;
}

int ensure_slot(vg_scene* s, uint32_t slot) {
  vg_ctx* ctx = s->ctx;
  if (!s->d_draws[slot].ptr) {
    if (int rc = ensure(s, s->d_draws[slot], sizeof(uint4) * 2 * static_cast<size_t>(ctx->cfg.max_meshlets))) return rc;
    VG_HIP(ctx, hipMemsetAsync(s->d_draws[slot].ptr, 0, s->d_draws[slot].bytes, ctx->stream));
    s->h_draw_ptrs[slot] = static_cast<uint4*>(s->d_draws[slot].ptr);
    VG_HIP(ctx, hipMemcpyAsync(static_cast<uint4**>(s->d_draw_ptrs.ptr) + slot, &s->h_draw_ptrs[slot], sizeof(uint4*),
                               hipMemcpyHostToDevice, ctx->stream));
    VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (ctx->cfg.expand && !s->d_indices[2 * slot].ptr) {
    // cameraIndex / shadowIndex (GeometryPool.ts:642-644): triangles * 3 indices per cut.
    const size_t caps[2] = {static_cast<size_t>(ctx->cfg.max_triangles) * 3, static_cast<size_t>(ctx->cfg.max_shadow_triangles) * 3};
    uint32_t* ptrs[2];
    for (int c = 0; c < 2; c++) {
      DevBuf& b = s->d_indices[2 * slot + c];
      if (int rc = ensure(s, b, sizeof(uint32_t) * caps[c])) return rc;
      VG_HIP(ctx, hipMemsetAsync(b.ptr, 0, b.bytes, ctx->stream));
      ptrs[c] = static_cast<uint32_t*>(b.ptr);
    }
    VG_HIP(ctx, hipMemcpyAsync(static_cast<uint32_t**>(s->d_index_ptrs.ptr) + 2 * slot, ptrs, sizeof(ptrs), hipMemcpyHostToDevice,
                               ctx->stream));
    VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return 0;
}

// Buffers of two-phase occlusion, sized for every view slot; the instance bitmap grows with the scene.
int ensure_two_phase(vg_scene* s) {
  vg_ctx* ctx = s->ctx;
  const size_t views = ctx->cfg.max_views;
  if (!s->d_phase.ptr) {
    if (int rc = ensure(s, s->d_counters2, sizeof(uint32_t) * s->counter_stride * views)) return rc;
    if (int rc = ensure(s, s->d_args2, sizeof(uint32_t) * s->args_stride * views)) return rc;
    if (int rc = ensure(s, s->d_reject_meshlets, sizeof(uint4) * ctx->cfg.max_meshlets * views)) return rc;
    if (int rc = ensure(s, s->d_phase, sizeof(uint32_t) * kPhaseWords * views)) return rc;
    VG_HIP(ctx, hipMemsetAsync(s->d_counters2.ptr, 0, s->d_counters2.bytes, ctx->stream));
    VG_HIP(ctx, hipMemsetAsync(s->d_args2.ptr, 0, s->d_args2.bytes, ctx->stream));
    VG_HIP(ctx, hipMemsetAsync(s->d_phase.ptr, 0, s->d_phase.bytes, ctx->stream));
  }
  const uint32_t words = static_cast<uint32_t>((s->i_col[0].size() + 31) / 32 + 1);
  if (s->reject_words < words) {
    if (int rc = ensure(s, s->d_reject_bits, sizeof(uint32_t) * words * views)) return rc;
    VG_HIP(ctx, hipMemsetAsync(s->d_reject_bits.ptr, 0, s->d_reject_bits.bytes, ctx->stream));
    s->reject_words = words;
    std::fill(s->phase1_ready.begin(), s->phase1_ready.end(), 0);
  }
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

Params make_params(vg_scene* s, Lane* lane, uint32_t view_count) {
  vg_ctx* ctx = s->ctx;
  Params p;
  std::memset(&p, 0, sizeof(p));
  p.mesh_table = static_cast<const float4*>(s->d_mesh_table.ptr);
  p.m_cull = static_cast<const float4*>(s->d_cull.ptr);
  p.m_lod = static_cast<const float4*>(s->d_lod.ptr);
  p.m_parent = static_cast<const float4*>(s->d_parent.ptr);
  p.m_err = static_cast<const float2*>(s->d_err.ptr);
  p.m_tris = static_cast<const uint32_t*>(s->d_tris.ptr);
  p.m_run = static_cast<const uint32_t*>(s->d_run.ptr);
  p.r_span = static_cast<const uint2*>(s->d_rspan.ptr);
  p.r_block = static_cast<const uint32_t*>(s->d_rblock.ptr);
  p.level_blocks = static_cast<const uint2*>(s->d_level_blocks.ptr);
  p.b_span = static_cast<const uint2*>(s->d_bspan.ptr);
  p.b_lod = static_cast<const float4*>(s->d_blod.ptr);
  p.b_par = static_cast<const float4*>(s->d_bpar.ptr);
  p.b_err = static_cast<const float4*>(s->d_berr.ptr);
  p.r_lod = static_cast<const float4*>(s->d_rlod.ptr);
  p.r_parent = static_cast<const float4*>(s->d_rparent.ptr);
  p.r_err = static_cast<const float2*>(s->d_rerr.ptr);
  p.i_col0 = static_cast<const float4*>(s->d_col[0].ptr);
  p.i_col1 = static_cast<const float4*>(s->d_col[1].ptr);
  p.i_col2 = static_cast<const float4*>(s->d_col[2].ptr);
  p.i_col3 = static_cast<const float4*>(s->d_col[3].ptr);
  p.c_sphere = static_cast<const float4*>(s->d_csphere.ptr);
  p.c_rec = static_cast<const float4*>(s->d_crec.ptr);
  p.c_bound = static_cast<const float4*>(s->d_cbound.ptr);
  p.cell_count = static_cast<uint32_t>(s->c_sphere.size());
  p.mesh_slots = s->mesh_slots;
  p.slots = s->slots;
  p.max_meshlets = ctx->cfg.max_meshlets;
  p.cut_capacity[0] = ctx->cfg.max_triangles * 3;
  p.cut_capacity[1] = ctx->cfg.max_shadow_triangles * 3;
  p.counters = static_cast<uint32_t*>(s->d_counters.ptr);
  p.counter_stride = s->counter_stride;
  p.args = static_cast<uint32_t*>(s->d_args.ptr);
  p.args_stride = s->args_stride;
  p.draws = static_cast<uint4* const*>(s->d_draw_ptrs.ptr);
  p.hzbs = static_cast<float* const*>(s->d_hzb_ptrs.ptr);
  p.metrics = static_cast<unsigned long long*>(s->d_metrics.ptr);
  p.views = static_cast<ViewDev*>(lane->d_views.ptr);
  p.view_count = view_count;
  p.cell_queue = static_cast<uint2*>(lane->d_cell_queue.ptr);
  p.cell_queue_count = static_cast<uint32_t*>(lane->d_cell_queue_count.ptr);
  p.work = static_cast<uint4*>(lane->d_work.ptr);
  p.work_start = static_cast<unsigned long long*>(lane->d_work_start.ptr);
  p.work_scale = static_cast<float*>(lane->d_work_scale.ptr);
  p.work_counter = static_cast<unsigned long long*>(lane->d_work_counter.ptr);
  p.work_capacity = lane->work_capacity / kWorkParts;  // per partition
  p.flags = static_cast<uint32_t*>(lane->d_flags.ptr);
  p.budget_thresholds = static_cast<const float*>(lane->d_budget_thresholds.ptr);
  p.budget_hist = static_cast<unsigned long long*>(lane->d_budget_hist.ptr);
  p.budget_bins = ctx->cfg.budget_bins;
  p.budget_block_hist = view_count * (ctx->cfg.budget_bins + 1) <= kBudgetBlockHistWords ? 1u : 0u;
  p.i_slot = static_cast<const uint32_t*>(s->d_islot.ptr);
  p.m_first = static_cast<const uint32_t*>(s->d_first.ptr);
  p.m_tri = static_cast<const uint32_t*>(s->d_tri.ptr);
  p.indices = static_cast<uint32_t* const*>(s->d_index_ptrs.ptr);
  p.expand_plan = static_cast<unsigned long long*>(lane->d_expand_plan.ptr);
  p.expand_records = static_cast<uint4*>(lane->d_expand_records.ptr);
  p.phase = 0;
  p.reject_bits = static_cast<uint32_t*>(s->d_reject_bits.ptr);
  p.reject_words = s->reject_words;
  p.reject_meshlets = static_cast<uint4*>(s->d_reject_meshlets.ptr);
  p.reject_capacity = ctx->cfg.max_meshlets;
  p.phase_words = static_cast<uint32_t*>(s->d_phase.ptr);
  p.counters1 = nullptr;
  if (s->footprint) {
    p.fp_instances = static_cast<uint32_t*>(s->d_fp_instances.ptr);
    p.fp_meshlets = static_cast<uint32_t*>(s->d_fp_meshlets.ptr);
  }
  return p;
}

constexpr int kKernelSlots = 13;
constexpr int kRasterSlot = 12;  // vg_hzb_from_cut's raster kernels, outside the cut's launch sequence
const char* const kKernelNames[kKernelSlots] = {"k_begin", "k_cells (+ budget passes)", "k_instances", "k_meshlets", "k_prefix(check 1)",
                                                "k_instances(retry 1)", "k_meshlets(retry 1)", "k_prefix(check 2)",
                                                "k_instances(retry 2)", "k_meshlets(retry 2)", "k_prefix(steer)",
                                                "k_expand (+ plan, resolve)", "k_raster (+ chunks)"};

hipEvent_t take_event(vg_scene* s) {
  if (!s->event_pool.empty()) {
    hipEvent_t e = s->event_pool.back();
    s->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// The launch sequence of one cut (GeometryPool.ts:1081-1092): begin, cells, instances, meshlets,
// prefix + check, 2 retry rounds (instances, meshlets, prefix + check / steer), then the expand pass on
// contexts that have it.
using MeshletKernel = void (*)(Params);
MeshletKernel meshlet_kernel(uint32_t waves_per_simd, bool affine = false) {
  switch (waves_per_simd) {
    case 4: return affine ? k_meshlets<4, true> : k_meshlets<4, false>;
    case 5: return affine ? k_meshlets<5, true> : k_meshlets<5, false>;
    case 6: return affine ? k_meshlets<6, true> : k_meshlets<6, false>;
    case 7: return affine ? k_meshlets<7, true> : k_meshlets<7, false>;
    default: return nullptr;
  }
}

bool affine_scene(const vg_scene* s) { return s->affine_instances && s->finite_runs; }

static_assert(kMaxMeshletsPerScene <= (1u << 24) && kMaxInstancesPerScene <= (1u << 24),
              "the meshlet pass's ring packs instance and meshlet ids in 24 bits");

bool schedule_valid(const vg_schedule& sc) {
  return sc.instance_blocks_per_cu >= 1 && sc.instance_blocks_per_cu <= 16 && sc.meshlet_blocks_per_cu >= 1 &&
         sc.meshlet_blocks_per_cu <= 32 && meshlet_kernel(sc.meshlet_waves_per_simd) != nullptr;
}

uint64_t schedule_key(const vg_schedule& sc) {
  return (static_cast<uint64_t>(sc.instance_blocks_per_cu) << 32) | (sc.meshlet_blocks_per_cu << 8) | sc.meshlet_waves_per_simd;
}

int enqueue_cut(vg_scene* s, Lane* lane, uint32_t view_count, bool capturing) {
  vg_ctx* ctx = s->ctx;
  hipStream_t st = lane->stream;
  Params p = make_params(s, lane, view_count);
  const uint32_t cap = lane->max_blocks_per_cu ? lane->max_blocks_per_cu : UINT32_MAX;
  const uint32_t persistent_instances = static_cast<uint32_t>(ctx->cus) * std::min(s->schedule.instance_blocks_per_cu, cap);
  const uint32_t persistent_meshlets = static_cast<uint32_t>(ctx->cus) * std::min(s->schedule.meshlet_blocks_per_cu, cap);
  const MeshletKernel k_meshlets = meshlet_kernel(s->schedule.meshlet_waves_per_simd, affine_scene(s));
  std::vector<hipEvent_t> ev;
  const bool timed = s->profile && !capturing && lane == &s->lane;
  if (timed) {
    for (int i = 0; i < kKernelSlots; i++) ev.push_back(take_event(s));
    ev.push_back(nullptr);  // the raster's slot is timed by vg_hzb_from_cut
  }
  int k = 0;
  auto mark = [&]() {
    if (timed && ev[k]) (void)hipEventRecord(ev[k], st);
    k++;
  };
  mark();
  hipLaunchKernelGGL(k_begin, dim3(view_count), dim3(256), 0, st, p);
  mark();
  if (p.cell_count) hipLaunchKernelGGL(k_cells, dim3((p.cell_count + 63) / 64, (view_count + 63) / 64), dim3(256), 0, st, p);
  if (lane->launch_budget) {
    hipLaunchKernelGGL(k_budget_instances, dim3(persistent_instances), dim3(256), 0, st, p);
    void (*const budget_meshlets)(Params) = affine_scene(s) ? k_budget_meshlets<true> : k_budget_meshlets<false>;
    // Dynamic LDS: one 32-bit histogram per launched view for the block, or a 64-bit one per wave.
    const size_t lds = p.budget_block_hist ? sizeof(uint32_t) * view_count * (p.budget_bins + 1)
                                           : sizeof(unsigned long long) * 4 * (p.budget_bins + 1);
    hipLaunchKernelGGL(budget_meshlets, dim3(persistent_meshlets), dim3(256), lds, st, p);
    hipLaunchKernelGGL(k_budget_pick, dim3(view_count), dim3(64), 0, st, p);
  }
  mark();
  hipLaunchKernelGGL(k_instances, dim3(persistent_instances), dim3(256), 0, st, p, 0u);
  mark();
  hipLaunchKernelGGL(k_meshlets, dim3(persistent_meshlets), dim3(256), 0, st, p);
  mark();
  hipLaunchKernelGGL(k_prefix, dim3(view_count), dim3(128), 0, st, p, 0u, 1u);
  mark();
  hipLaunchKernelGGL(k_instances, dim3(persistent_instances), dim3(256), 0, st, p, 1u);
  mark();
  hipLaunchKernelGGL(k_meshlets, dim3(persistent_meshlets), dim3(256), 0, st, p);
  mark();
  hipLaunchKernelGGL(k_prefix, dim3(view_count), dim3(128), 0, st, p, 0u, 2u);
  mark();
  hipLaunchKernelGGL(k_instances, dim3(persistent_instances), dim3(256), 0, st, p, 2u);
  mark();
  hipLaunchKernelGGL(k_meshlets, dim3(persistent_meshlets), dim3(256), 0, st, p);
  mark();
  hipLaunchKernelGGL(k_prefix, dim3(view_count), dim3(128), 0, st, p, 1u, 0u);
  mark();
  if (ctx->cfg.expand) {
    hipLaunchKernelGGL(k_expand_plan, dim3(1), dim3(64), 0, st, p);
    hipLaunchKernelGGL(k_expand_resolve, dim3(static_cast<uint32_t>(ctx->cus) * kExpandBlocksPerCu), dim3(256), 0, st, p);
    hipLaunchKernelGGL(k_expand, dim3(static_cast<uint32_t>(ctx->cus) * kExpandBlocksPerCu), dim3(256), 0, st, p);
  }
  mark();
  if (timed) s->pending.push_back(ev);
  VG_HIP(ctx, hipGetLastError());
  return 0;
}

int enqueue_phase2(vg_scene* s, Lane* lane, uint32_t view_count) {
  vg_ctx* ctx = s->ctx;
  hipStream_t st = lane->stream;
  Params p = make_params(s, lane, view_count);
  p.phase = 2;
  p.counters1 = static_cast<const uint32_t*>(s->d_counters.ptr);
  p.counters = static_cast<uint32_t*>(s->d_counters2.ptr);
  p.args = static_cast<uint32_t*>(s->d_args2.ptr);
  const uint32_t cap = lane->max_blocks_per_cu ? lane->max_blocks_per_cu : UINT32_MAX;
  const uint32_t persistent_instances = static_cast<uint32_t>(ctx->cus) * std::min(s->schedule.instance_blocks_per_cu, cap);
  const uint32_t persistent_meshlets = static_cast<uint32_t>(ctx->cus) * std::min(s->schedule.meshlet_blocks_per_cu, cap);
  const MeshletKernel k_meshlets = meshlet_kernel(s->schedule.meshlet_waves_per_simd, affine_scene(s));
  hipLaunchKernelGGL(k_begin, dim3(view_count), dim3(256), 0, st, p);
  if (p.cell_count) hipLaunchKernelGGL(k_cells, dim3((p.cell_count + 63) / 64, (view_count + 63) / 64), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_instances, dim3(persistent_instances), dim3(256), 0, st, p, 0u);
  hipLaunchKernelGGL(k_meshlets, dim3(persistent_meshlets), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_retest_plan, dim3(1), dim3(64), 0, st, p);
  hipLaunchKernelGGL(k_retest, dim3(static_cast<uint32_t>(ctx->cus) * kExpandBlocksPerCu), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_prefix, dim3(view_count), dim3(128), 0, st, p, 2u, 0u);
  if (ctx->cfg.expand) {
    hipLaunchKernelGGL(k_expand_plan, dim3(1), dim3(64), 0, st, p);
    hipLaunchKernelGGL(k_expand_resolve, dim3(static_cast<uint32_t>(ctx->cus) * kExpandBlocksPerCu), dim3(256), 0, st, p);
    hipLaunchKernelGGL(k_expand, dim3(static_cast<uint32_t>(ctx->cus) * kExpandBlocksPerCu), dim3(256), 0, st, p);
  }
  VG_HIP(ctx, hipGetLastError());
  return 0;
}

int lane_reserve_views(vg_ctx* ctx, Lane* lane) {
  VG_HIP(ctx, hipStreamSynchronize(lane->stream));
  if (int rc = ensure_lane(ctx, lane, lane->d_views, sizeof(ViewDev) * ctx->cfg.max_views)) return rc;
  // chunk plans of the expand pass and of the second phase's re-test
  if (int rc = ensure_lane(ctx, lane, lane->d_expand_plan, sizeof(unsigned long long) * (2 * static_cast<size_t>(ctx->cfg.max_views) + 1)))
    return rc;
  if (ctx->cfg.expand) {
    const size_t chunks = 2 * static_cast<size_t>(ctx->cfg.max_views) * ((ctx->cfg.max_meshlets + 63) / 64);
    if (int rc = ensure_lane(ctx, lane, lane->d_expand_records, sizeof(uint4) * 64 * chunks)) return rc;
  }
  for (int r = 0; r < Lane::kRing; r++) {
    if (lane->h_ring[r]) (void)hipHostFree(lane->h_ring[r]);
    if (lane->h_ring_thresholds[r]) (void)hipHostFree(lane->h_ring_thresholds[r]);
    lane->h_ring[r] = nullptr;
    lane->h_ring_thresholds[r] = nullptr;
    lane->ring_pending[r] = false;
    VG_HIP(ctx, hipHostMalloc(&lane->h_ring[r], sizeof(ViewDev) * ctx->cfg.max_views, hipHostMallocDefault));
    VG_HIP(ctx, hipHostMalloc(reinterpret_cast<void**>(&lane->h_ring_thresholds[r]),
                              sizeof(float) * ctx->cfg.max_views * ctx->cfg.budget_bins, hipHostMallocDefault));
  }
  if (int rc = ensure_lane(ctx, lane, lane->d_budget_thresholds, sizeof(float) * ctx->cfg.max_views * ctx->cfg.budget_bins)) return rc;
  if (int rc = ensure_lane(ctx, lane, lane->d_budget_hist, sizeof(unsigned long long) * ctx->cfg.max_views * (ctx->cfg.budget_bins + 1)))
    return rc;
  lane->view_capacity = ctx->cfg.max_views;
  return 0;
}

// The device view records of `views` (vg_view -> ViewDev) and, for views with a triangle budget, their
// threshold tables (budget_bins floats per view). Returns whether any view has a budget.
bool stage_views(vg_scene* s, const vg_view* views, uint32_t count, bool fresh, ViewDev* staging, float* thresholds) {
  vg_ctx* ctx = s->ctx;
  bool budget = false;
  for (uint32_t i = 0; i < count; i++) {
    const vg_view& in = views[i];
    ViewDev& v = staging[i];
    std::memset(&v, 0, sizeof(v));
    std::memcpy(v.view, in.view_matrix, sizeof(v.view));
    std::memcpy(v.proj, in.projection_matrix, sizeof(v.proj));
    std::memcpy(v.planes, in.frustum_planes, sizeof(v.planes));
    std::memcpy(v.sweep, in.shadow_sweep, sizeof(float) * 3);
    v.proj_scale = in.proj_scale;
    v.near_plane = in.near_plane;
    v.error_threshold = in.error_threshold;
    v.shadow_error_scale = in.shadow_error_scale;
    v.lod_blend_max = in.lod_blend_max;
    v.lod_scale_limit = in.lod_scale_limit;
    v.min_pixel_radius = in.min_pixel_radius;
    v.max_error_threshold = in.max_error_threshold;
    v.frustum = in.frustum_culling;
    v.occlusion = in.occlusion;
    v.slot = in.slot;
    v.budget = in.triangle_budget;
    v.range_lo = in.threshold_range[0];
    v.range_hi = in.threshold_range[1];
    v.hzb_valid = (in.occlusion && s->hzb_valid[in.slot]) ? 1u : 0u;
    v.fresh = fresh ? 1u : 0u;
    v.two_phase = in.occlusion == 2 ? 1u : 0u;
    if (in.triangle_budget) {
      budget = true;
      vg_budget_thresholds(in.threshold_range[0], in.threshold_range[1], ctx->cfg.budget_bins,
                           thresholds + static_cast<size_t>(i) * ctx->cfg.budget_bins);
    }
  }
  return budget;
}

// The checks every cut entry point makes on its views.
int check_views(vg_ctx* ctx, const vg_view* views, uint32_t count, const char* who) {
  const std::string w(who);
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, w + ": host-only context (device < 0)");
  if (!views || count == 0) return fail(ctx, VG_E_INVALID, w + ": no views");
  if (count > ctx->cfg.max_views || count > kMaxLaunchViews) return fail(ctx, VG_E_INVALID, w + ": more views than max_views");
  uint64_t seen[kMaxLaunchViews / 64] = {0};
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t slot = views[i].slot;
    if (slot >= ctx->cfg.max_views) return fail(ctx, VG_E_INVALID, w + ": view slot out of range");
    if (seen[slot / 64] >> (slot % 64) & 1ull) return fail(ctx, VG_E_INVALID, w + ": two views share a slot");
    seen[slot / 64] |= 1ull << (slot % 64);
    if (views[i].triangle_budget && !(views[i].threshold_range[0] > 0 && views[i].threshold_range[1] >= views[i].threshold_range[0]))
      return fail(ctx, VG_E_INVALID, w + ": a triangle budget needs 0 < threshold_range[0] <= threshold_range[1]");
    if (!(views[i].proj_scale > 0) || !(views[i].near_plane > 0))
      return fail(ctx, VG_E_INVALID, w + ": proj_scale and near_plane must be positive");
    if (views[i].occlusion > 2) return fail(ctx, VG_E_INVALID, w + ": occlusion must be 0, 1 or 2");
  }
  return 0;
}

int prepare_cut(vg_scene* s, Lane* lane, const vg_view* views, uint32_t count, bool fresh) {
  vg_ctx* ctx = s->ctx;
  if (int rc = check_views(ctx, views, count, "vg_cut")) return rc;
  if (int rc = upload_scene(s)) return rc;
  for (uint32_t i = 0; i < count; i++)
    if (views[i].occlusion == 2) {
      if (int rc = ensure_two_phase(s)) return rc;
      break;
    }
  for (uint32_t i = 0; i < count; i++)
    if (int rc = ensure_slot(s, views[i].slot)) return rc;
  size_t queue_bytes = 0, want = 0;
  transient_sizes(s, count, &queue_bytes, &want);
  if (lane->borrowed) {
    // The caller's work list may be smaller than the worst case: an overflow sets the capacity flag
    // (the launch then reports VG_E_CAPACITY). The cell queue must hold every (view, cell).
    if (lane->d_cell_queue.bytes < queue_bytes || lane->work_capacity == 0)
      return fail(ctx, VG_E_INVALID, "vg_cut: lane buffers smaller than the launch needs");
  } else {
    if (lane->cell_queue_capacity < queue_bytes / sizeof(uint2)) {
      if (int rc = ensure_lane(ctx, lane, lane->d_cell_queue, queue_bytes)) return rc;
      lane->cell_queue_capacity = queue_bytes / sizeof(uint2);
    }
    if (lane->work_capacity < want) {
      if (int rc = ensure_lane(ctx, lane, lane->d_work, sizeof(uint4) * want)) return rc;
      if (int rc = ensure_lane(ctx, lane, lane->d_work_start, sizeof(unsigned long long) * want)) return rc;
      if (int rc = ensure_lane(ctx, lane, lane->d_work_scale, sizeof(float) * want)) return rc;
      lane->work_capacity = static_cast<uint32_t>(want);
    }
  }
  if (lane->view_capacity < count)
    if (int rc = lane_reserve_views(ctx, lane)) return rc;
  const int ring = static_cast<int>(lane->ring_next++ % Lane::kRing);
  if (lane->ring_pending[ring]) {
    VG_HIP(ctx, hipEventSynchronize(lane->ring_copied[ring]));
    lane->ring_pending[ring] = false;
  }
  ViewDev* staging = static_cast<ViewDev*>(lane->h_ring[ring]);
  lane->launch_budget = stage_views(s, views, count, fresh, staging, lane->h_ring_thresholds[ring]);
  if (lane->launch_budget)
    VG_HIP(ctx, hipMemcpyAsync(lane->d_budget_thresholds.ptr, lane->h_ring_thresholds[ring], sizeof(float) * count * ctx->cfg.budget_bins,
                               hipMemcpyHostToDevice, lane->stream));
  VG_HIP(ctx, hipMemcpyAsync(lane->d_views.ptr, staging, sizeof(ViewDev) * count, hipMemcpyHostToDevice, lane->stream));
  VG_HIP(ctx, hipEventRecord(lane->ring_copied[ring], lane->stream));
  lane->ring_pending[ring] = true;
  return 0;
}

int launch_graph(vg_scene* s, Lane* lane, uint32_t count) {
  vg_ctx* ctx = s->ctx;
  // Everything enqueue_cut bakes into the captured launches: buffer addresses (the layout versions count
  // reallocations), and the scalars and grids taken from the scene's contents (cell count, used mesh
  // slots, the reject bitmap's words), which change without a reallocation when instances are set.
  const uint64_t key[8] = {s->layout_version, lane->layout_version, static_cast<uint64_t>(lane->launch_budget), schedule_key(s->schedule),
                           affine_scene(s) ? 1u : 0u, static_cast<uint64_t>(s->c_sphere.size()),
                           (static_cast<uint64_t>(s->mesh_slots) << 32) | s->reject_words, static_cast<uint64_t>(s->footprint)};
  if (!lane->graph_exec || lane->graph_views != count || std::memcmp(key, lane->graph_key, sizeof(key)) != 0) {
    if (lane->graph_exec) (void)hipGraphExecDestroy(lane->graph_exec);
    if (lane->graph) (void)hipGraphDestroy(lane->graph);
    lane->graph_exec = nullptr;
    lane->graph = nullptr;
    VG_HIP(ctx, hipStreamBeginCapture(lane->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_cut(s, lane, count, true);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(lane->stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
    lane->graph = graph;
    VG_HIP(ctx, hipGraphInstantiate(&lane->graph_exec, graph, nullptr, nullptr, 0));
    lane->graph_views = count;
    std::memcpy(lane->graph_key, key, sizeof(key));
  }
  VG_HIP(ctx, hipGraphLaunch(lane->graph_exec, lane->stream));
  return VG_OK;
}

int read_slot(vg_scene* s, hipStream_t stream, uint32_t view_slot, uint32_t* counters, uint32_t* draw_args, uint32_t* draw_list,
              uint32_t max_entries, vg_metrics* metrics) {
  vg_ctx* ctx = s->ctx;
  if (counters)
    VG_HIP(ctx, hipMemcpyAsync(counters, static_cast<uint32_t*>(s->d_counters.ptr) + static_cast<size_t>(view_slot) * s->counter_stride,
                               sizeof(uint32_t) * s->counter_stride, hipMemcpyDeviceToHost, stream));
  if (draw_args)
    VG_HIP(ctx, hipMemcpyAsync(draw_args, static_cast<uint32_t*>(s->d_args.ptr) + static_cast<size_t>(view_slot) * s->args_stride,
                               sizeof(uint32_t) * s->args_stride, hipMemcpyDeviceToHost, stream));
  if (draw_list && max_entries) {
    const uint32_t n = std::min(max_entries, ctx->cfg.max_meshlets);
    if (!s->d_draws[view_slot].ptr) {
      std::memset(draw_list, 0, sizeof(uint4) * 2 * static_cast<size_t>(max_entries));
    } else {
      const uint4* d = static_cast<const uint4*>(s->d_draws[view_slot].ptr);
      VG_HIP(ctx, hipMemcpyAsync(draw_list, d, sizeof(uint4) * n, hipMemcpyDeviceToHost, stream));
      VG_HIP(ctx, hipMemcpyAsync(draw_list + static_cast<size_t>(max_entries) * 4, d + ctx->cfg.max_meshlets, sizeof(uint4) * n,
                                 hipMemcpyDeviceToHost, stream));
    }
  }
  unsigned long long m[kMetricWords];
  if (metrics)
    VG_HIP(ctx, hipMemcpyAsync(m, static_cast<unsigned long long*>(s->d_metrics.ptr) + static_cast<size_t>(view_slot) * kMetricWords,
                               sizeof(m), hipMemcpyDeviceToHost, stream));
  VG_HIP(ctx, hipStreamSynchronize(stream));
  if (metrics) {
    std::memset(metrics, 0, sizeof(*metrics));
    metrics->cells_tested = m[M_CELLS];
    metrics->instances_tested = m[M_INSTANCES];
    metrics->work_items[0] = m[M_WORK0];
    metrics->work_items[1] = m[M_WORK1];
    metrics->clusters_tested[0] = m[M_CLUSTERS0];
    metrics->clusters_tested[1] = m[M_CLUSTERS1];
    metrics->occlusion_tests = m[M_OCCLUSION];
    float t;
    const uint32_t bits = static_cast<uint32_t>(m[M_THRESHOLD]);
    std::memcpy(&t, &bits, 4);
    metrics->error_threshold = t;
    metrics->budget_bin = static_cast<uint32_t>(m[M_BUDGET_BIN]);
    metrics->budget_cluster_tests = m[M_BUDGET_TESTS];
  }
  return VG_OK;
}

}  // namespace vgi

// ------------------------------------------------------------------ C ABI

extern "C" {

int vg_create(const vg_config* config, vg_ctx** out) {
  if (!out) return VG_E_INVALID;
  *out = nullptr;
  vg_config c;
  if (config) {
    c = *config;
  } else {
    std::memset(&c, 0, sizeof(c));
  }
  if (!c.max_triangles) c.max_triangles = 10000000;
  if (!c.max_shadow_triangles) c.max_shadow_triangles = 4000000;
  if (!c.max_meshlets) c.max_meshlets = 1u << 20;
  if (!c.max_views) c.max_views = 1;
  if (!c.mesh_slots) c.mesh_slots = 64;
  if (!c.hzb_max_size) c.hzb_max_size = 1024;
  if (!c.budget_bins) c.budget_bins = 256;
  if (c.budget_bins > kMaxBudgetBins || c.max_meshlets > (1u << 24) || c.mesh_slots > 65535 || c.max_views > kMaxLaunchViews || c.hzb_max_size > 4096 ||
      c.max_triangles > 0x55555555u || c.max_shadow_triangles > 0x55555555u || c.expand > 1)
    return VG_E_INVALID;
  auto* ctx = new (std::nothrow) vg_ctx();
  if (!ctx) return VG_E_NOMEM;
  ctx->cfg = c;
  if (c.device < 0) {  // host-only: scenes lay out their buffers, nothing runs
    *out = ctx;
    return VG_OK;
  }
  if (hipSetDevice(c.device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return VG_E_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c.device) == hipSuccess && prop.multiProcessorCount > 0) ctx->cus = prop.multiProcessorCount;
  // vg_cut_static's scratch: this device's copy of the array, and host staging sized for its largest call.
  void* scratch_ptr = nullptr;
  if (hipGetSymbolAddress(&scratch_ptr, HIP_SYMBOL(scratch)) != hipSuccess || !scratch_ptr) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return VG_E_DEVICE;
  }
  ctx->static_scratch = scratch_ptr;
  ctx->static_views.assign(sizeof(ViewDev) * kStaticViews, 0);
  ctx->static_thresholds.assign(static_cast<size_t>(kStaticViews) * c.budget_bins, 0.0f);
  *out = ctx;
  return VG_OK;
}

void vg_destroy(vg_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

const char* vg_last_error(const vg_ctx* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

uint32_t vg_counter_words(const vg_ctx* ctx) { return state_base(ctx->cfg.mesh_slots) + kStateWords; }
uint32_t vg_draw_args_words(const vg_ctx* ctx) { return ctx->cfg.mesh_slots * 20; }

int vg_scene_create(vg_ctx* ctx, vg_scene** out) {
  if (!ctx || !out) return VG_E_INVALID;
  *out = nullptr;
  const bool device = ctx->cfg.device >= 0;
  if (device) VG_HIP(ctx, hipSetDevice(ctx->cfg.device));
  auto* s = new (std::nothrow) vg_scene();
  if (!s) return fail(ctx, VG_E_NOMEM, "out of host memory");
  s->ctx = ctx;
  s->slots = ctx->cfg.mesh_slots;
  s->meshes.resize(s->slots);
  s->mesh_table.assign(static_cast<size_t>(s->slots) * kMeshStride, make_float4(0, 0, 0, 0));
  const uint32_t views = ctx->cfg.max_views;
  s->counter_stride = state_base(s->slots) + kStateWords;
  s->args_stride = s->slots * 20;
  s->d_draws.resize(views);
  s->d_hzbs.resize(views);
  s->d_indices.resize(2 * static_cast<size_t>(views));
  s->phase1_ready.assign(views, 0);
  s->phase2_ran.assign(views, 0);
  s->hzb_valid.assign(views, 0);
  s->h_draw_ptrs.assign(views, nullptr);
  s->h_hzb_ptrs.assign(views, nullptr);
  int rc = 0;
  auto init = [&]() -> int {
    if ((rc = ensure(s, s->d_counters, sizeof(uint32_t) * s->counter_stride * views)) ||
        (rc = ensure(s, s->d_args, sizeof(uint32_t) * s->args_stride * views)) ||
        (rc = ensure(s, s->d_metrics, sizeof(unsigned long long) * kMetricWords * views)) ||
        (rc = ensure(s, s->d_draw_ptrs, sizeof(uint4*) * views)) || (rc = ensure(s, s->d_hzb_ptrs, sizeof(float*) * views)) ||
        (rc = ensure(s, s->d_index_ptrs, sizeof(uint32_t*) * 2 * views)) || (rc = lane_init(ctx, &s->lane, ctx->stream)))
      return rc;
    // Controller state: scale 1 for both cuts (GeometryPool.ts:602-603), everything else 0.
    std::vector<uint32_t> counters(static_cast<size_t>(s->counter_stride) * views, 0u);
    for (uint32_t v = 0; v < views; v++) {
      counters[static_cast<size_t>(v) * s->counter_stride + state_base(s->slots) + kStateScale] = 10000;
      counters[static_cast<size_t>(v) * s->counter_stride + state_base(s->slots) + kStateScale + 1] = 10000;
    }
    VG_HIP(ctx, hipMemcpy(s->d_counters.ptr, counters.data(), sizeof(uint32_t) * counters.size(), hipMemcpyHostToDevice));
    VG_HIP(ctx, hipMemset(s->d_args.ptr, 0, s->d_args.bytes));
    VG_HIP(ctx, hipMemset(s->d_metrics.ptr, 0, s->d_metrics.bytes));
    VG_HIP(ctx, hipMemset(s->d_draw_ptrs.ptr, 0, s->d_draw_ptrs.bytes));
    VG_HIP(ctx, hipMemset(s->d_hzb_ptrs.ptr, 0, s->d_hzb_ptrs.bytes));
    VG_HIP(ctx, hipMemset(s->d_index_ptrs.ptr, 0, s->d_index_ptrs.bytes));
    return 0;
  };
  if (device && (rc = init())) {
    vg_scene_destroy(s);
    return rc;
  }
  *out = s;
  return VG_OK;
}

void vg_scene_destroy(vg_scene* s) {
  if (!s) return;
  if (s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
  lane_destroy(&s->lane);
  DevBuf* bufs[] = {&s->d_run, &s->d_rspan, &s->d_rlod, &s->d_rparent, &s->d_rerr, &s->d_rblock, &s->d_level_blocks, &s->d_bspan, &s->d_blod, &s->d_bpar,
                    &s->d_berr, &s->d_mesh_table, &s->d_cull, &s->d_lod, &s->d_parent, &s->d_err, &s->d_tris, &s->d_col[0], &s->d_col[1],
                    &s->d_col[2], &s->d_col[3], &s->d_csphere, &s->d_crec, &s->d_cbound, &s->d_counters, &s->d_args, &s->d_metrics,
                    &s->d_draw_ptrs, &s->d_hzb_ptrs, &s->d_hzb_stage, &s->d_fp_instances, &s->d_fp_meshlets, &s->d_fp_counts,
                    &s->d_first, &s->d_tri, &s->d_islot, &s->d_index_ptrs, &s->d_counters2, &s->d_args2, &s->d_reject_bits,
                    &s->d_reject_meshlets, &s->d_phase, &s->d_vpos, &s->d_vlist, &s->d_raster_chunks,
                    &s->d_raster_count, &s->d_raster_views};
  for (DevBuf* b : bufs) release(s, *b);
  for (DevBuf& b : s->d_draws) release(s, b);
  for (DevBuf& b : s->d_indices) release(s, b);
  for (DevBuf& b : s->d_hzbs) release(s, b);
  for (auto& seq : s->pending)
    for (hipEvent_t e : seq)
      if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  delete s;
}

uint64_t vg_scene_bytes(const vg_scene* s) { return s ? s->device_bytes + s->lane.device_bytes : 0; }

int vg_mesh_add(vg_scene* s, const vg_mesh_desc* mesh, uint32_t* slot_out) {
  if (!s || !mesh || !slot_out) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  const uint32_t n = mesh->meshlet_count, L = mesh->level_count;
  if (n == 0 || L == 0 || !mesh->meshlet_bounds || !mesh->meshlet_info || !mesh->level_ranges || !mesh->level_errors ||
      !mesh->level_guard_errors)
    return fail(ctx, VG_E_INVALID, "vg_mesh_add: empty or incomplete mesh");
  if (s->m_cull.size() + n > kMaxMeshletsPerScene) return fail(ctx, VG_E_CAPACITY, "vg_mesh_add: more than 2^24 meshlets in a scene");
  for (uint32_t i = 0; i < n; i++)
    if (mesh->meshlet_info[static_cast<size_t>(i) * 4] > 255)  // the draw list packs it into 8 bits (GeometryPool.ts:957-962)
      return fail(ctx, VG_E_INVALID, "vg_mesh_add: a meshlet has more than 255 triangles");
  if (ctx->cfg.expand && !mesh->meshlet_triangles)
    return fail(ctx, VG_E_INVALID, "vg_mesh_add: the expand pass needs the meshlet triangles");
  if (mesh->meshlet_triangles) {
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t* info = mesh->meshlet_info + static_cast<size_t>(i) * 4;
      if (info[1] % 3 != 0 || static_cast<uint64_t>(info[1] / 3) + info[0] > mesh->triangle_count)
        return fail(ctx, VG_E_INVALID, "vg_mesh_add: a meshlet's triangles lie outside meshlet_triangles");
    }
    if (s->index_used + mesh->triangle_count * 3 > 0xffffffffull)
      return fail(ctx, VG_E_CAPACITY, "vg_mesh_add: more than 2^32 indices in a scene");
  }
  // Occluder raster inputs: each meshlet's vertex list holds its largest local id (GeometryPool.ts:1102).
  const bool has_vertices = mesh->positions && mesh->meshlet_vertices && mesh->vertex_count && mesh->meshlet_vertex_count;
  std::vector<uint32_t> vcount;
  if (has_vertices) {
    if (!mesh->meshlet_triangles) return fail(ctx, VG_E_INVALID, "vg_mesh_add: positions need the meshlet triangles too");
    if (s->m_verts.size() + mesh->vertex_count > 0xffffffffull || s->m_mverts.size() + mesh->meshlet_vertex_count > 0xfffffffeull)
      return fail(ctx, VG_E_CAPACITY, "vg_mesh_add: more than 2^32 vertices in a scene");
    for (uint64_t i = 0; i < mesh->meshlet_vertex_count; i++)
      if (mesh->meshlet_vertices[i] >= mesh->vertex_count) return fail(ctx, VG_E_INVALID, "vg_mesh_add: a meshlet vertex id >= vertex_count");
    vcount.resize(n);
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t* info = mesh->meshlet_info + static_cast<size_t>(i) * 4;
      uint32_t top = 0;
      for (uint32_t t = 0; t < info[0]; t++) {
        const uint32_t packed = mesh->meshlet_triangles[info[1] / 3 + t];
        top = std::max(top, std::max(packed & 255u, std::max((packed >> 8) & 255u, (packed >> 16) & 255u)) + 1);
      }
      if (top > 128 || static_cast<uint64_t>(info[3]) + top > mesh->meshlet_vertex_count)
        return fail(ctx, VG_E_INVALID, "vg_mesh_add: a meshlet's local ids lie outside its vertex list (or above 127)");
      vcount[i] = top;
    }
  }
  uint32_t slot = 0;
  while (slot < s->slots && s->meshes[slot].used) slot++;
  if (slot >= s->slots) return fail(ctx, VG_E_CAPACITY, "vg_mesh_add: no free mesh slot");
  for (uint32_t l = 0; l < L; l++) {
    const uint64_t end = static_cast<uint64_t>(mesh->level_ranges[l * 2]) + mesh->level_ranges[l * 2 + 1];
    if (end > n) return fail(ctx, VG_E_INVALID, "vg_mesh_add: level range outside the meshlets");
  }
  MeshHost& m = s->meshes[slot];
  m = MeshHost();
  m.used = true;
  m.meshlet_base = static_cast<uint32_t>(s->m_cull.size());
  m.meshlet_count = n;
  std::memcpy(m.bounding, mesh->bounding_sphere, sizeof(m.bounding));
  std::memcpy(m.lod_bounds, mesh->lod_bounds_sphere, sizeof(m.lod_bounds));
  std::memcpy(m.cull_bounds, mesh->cull_bounds_sphere, sizeof(m.cull_bounds));
  for (uint32_t i = 0; i < n; i++) {
    const float* b = mesh->meshlet_bounds + static_cast<size_t>(i) * 16;
    s->m_lod.push_back(make_float4(b[0], b[1], b[2], b[3]));
    s->m_parent.push_back(make_float4(b[4], b[5], b[6], b[7]));
    s->m_err.push_back(make_float2(b[8], b[9]));
    s->m_cull.push_back(make_float4(b[12], b[13], b[14], b[15]));
    s->m_tris.push_back(mesh->meshlet_info[static_cast<size_t>(i) * 4]);
    const uint32_t* info = mesh->meshlet_info + static_cast<size_t>(i) * 4;
    s->m_info.push_back(info[0]);
    s->m_info.push_back(static_cast<uint32_t>(info[1] + s->index_used));
    s->m_info.push_back(info[2]);
    // vertexOffset rebased to the pool's meshlet vertex list (GeometryPool.ts:225, 428); no list: ~0
    const uint32_t voff = has_vertices ? static_cast<uint32_t>(info[3] + s->m_mverts.size()) : 0xffffffffu;
    s->m_info.push_back(voff);
    s->m_vlist.push_back(make_uint2(has_vertices ? voff : 0u, has_vertices ? vcount[i] : 0u));
    s->m_first.push_back(static_cast<uint32_t>(info[1] + s->index_used));
  }
  // LOD runs: consecutive meshlets of one level with the same LOD-test inputs (bit-equal spheres and
  // errors; any sphere with a zero error or an error >= 5e29, which project to 0 and 1e30).
  {
    auto bits = [](float f) {
      uint32_t u;
      std::memcpy(&u, &f, 4);
      return u;
    };
    const float inf_half = static_cast<float>(1e30 * 0.5);
    auto key_equal = [&](size_t a, size_t b, int which) {
      const float ea = which ? s->m_err[a].y : s->m_err[a].x, eb = which ? s->m_err[b].y : s->m_err[b].x;
      if (bits(ea) != bits(eb)) return false;
      if (ea == 0.0f || ea >= inf_half) return true;
      const float4 sa = which ? s->m_parent[a] : s->m_lod[a], sb = which ? s->m_parent[b] : s->m_lod[b];
      return bits(sa.x) == bits(sb.x) && bits(sa.y) == bits(sb.y) && bits(sa.z) == bits(sb.z) && bits(sa.w) == bits(sb.w);
    };
    std::vector<uint8_t> level_start(n, 0);
    for (uint32_t l = 0; l < L; l++) {
      const uint32_t start = mesh->level_ranges[l * 2], end = start + mesh->level_ranges[l * 2 + 1];
      if (start < end) level_start[start] = 1;
      if (start < end && end < n) level_start[end] = 1;
    }
    for (uint32_t i = 0; i < n;) {
      const size_t a = m.meshlet_base + i;
      uint32_t len = 1;
      while (i + len < n && len < kMaxRun && !level_start[i + len] && key_equal(a, a + len, 0) && key_equal(a, a + len, 1)) len++;
      const uint32_t run = static_cast<uint32_t>(s->r_span.size());
      s->r_span.push_back(make_uint2(static_cast<uint32_t>(a), len));
      s->r_lod.push_back(s->m_lod[a]);
      s->r_parent.push_back(s->m_parent[a]);
      s->r_err.push_back(s->m_err[a]);
      for (uint32_t k = 0; k < len; k++) s->m_run.push_back(run);
      // A run opens a new block at a level boundary or when the current block is full.
      if (i == 0 || level_start[i] || s->b_span.back().y == kBlockRuns) s->b_span.push_back(make_uint2(run, 0));
      s->b_span.back().y++;
      s->r_block.push_back(static_cast<uint32_t>(s->b_span.size() - 1));
      i += len;
    }
    // The new blocks' bounds.
    for (size_t b = s->b_lod.size(); b < s->b_span.size(); b++) {
      float4 lod, par, err;
      block_bounds(s->r_lod.data(), s->r_parent.data(), s->r_err.data(), s->b_span[b].x, s->b_span[b].y, &lod, &par, &err);
      s->b_lod.push_back(lod);
      s->b_par.push_back(par);
      s->b_err.push_back(err);
    }
  }
  // The pool's triangle array grows by the mesh's triangles (GeometryPool.ts:326, 430); firstIndex is
  // rebased by 3 x its base. A mesh given without triangles takes the space its meshlets' counts need.
  uint64_t mesh_indices = 0;
  for (uint32_t i = 0; i < n; i++) mesh_indices += static_cast<uint64_t>(mesh->meshlet_info[static_cast<size_t>(i) * 4]) * 3;
  if (mesh->meshlet_triangles) {
    s->m_tri.resize(s->index_used / 3, 0u);
    s->m_tri.insert(s->m_tri.end(), mesh->meshlet_triangles, mesh->meshlet_triangles + mesh->triangle_count);
    mesh_indices = mesh->triangle_count * 3;
  }
  s->index_used += mesh_indices;
  if (has_vertices) {
    const uint32_t vbase = static_cast<uint32_t>(s->m_verts.size());
    for (uint32_t v = 0; v < mesh->vertex_count; v++)
      s->m_verts.push_back(make_float4(mesh->positions[v * 3ull], mesh->positions[v * 3ull + 1], mesh->positions[v * 3ull + 2], 0.0f));
    for (uint64_t i = 0; i < mesh->meshlet_vertex_count; i++) s->m_mverts.push_back(mesh->meshlet_vertices[i] + vbase);
  }
  // Level table after the mesh records: (first meshlet, count, min own, max parent), (max own, min parent, 0, 0)
  // (GeometryPool.ts:432-445).
  m.level_count = L;
  m.level_base = s->slots * kMeshStride + s->level_used;
  s->level_used += L * 2;
  s->mesh_table.resize(static_cast<size_t>(s->slots) * kMeshStride + s->level_used);
  // Per level, the LOD blocks of its first meshlet and of its last one (the instance pass's work items cover
  // the blocks from the first possible level's first meshlet to the last possible level's last one).
  s->level_blocks.resize(s->level_used / 2);
  for (uint32_t l = 0; l < L; l++) {
    s->mesh_table[m.level_base + l * 2] =
        make_float4(static_cast<float>(mesh->level_ranges[l * 2] + m.meshlet_base), static_cast<float>(mesh->level_ranges[l * 2 + 1]),
                    mesh->level_errors[l * 2], mesh->level_errors[l * 2 + 1]);
    s->mesh_table[m.level_base + l * 2 + 1] = make_float4(mesh->level_guard_errors[l * 2], mesh->level_guard_errors[l * 2 + 1], 0, 0);
    const uint64_t first = static_cast<uint64_t>(mesh->level_ranges[l * 2]) + m.meshlet_base, end = first + mesh->level_ranges[l * 2 + 1];
    uint2 blocks = make_uint2(0, 0);
    if (first < s->m_run.size()) blocks.x = s->r_block[s->m_run[first]];
    if (end >= 1 && end - 1 < s->m_run.size()) blocks.y = s->r_block[s->m_run[end - 1]] + 1;
    s->level_blocks[(m.level_base - s->slots * kMeshStride) / 2 + l] = blocks;
  }
  write_mesh_record(s, slot);
  // writeMeshRecord defaults (GeometryPool.ts:350): no draw distance, min pixel radius 0, enabled, shadow, no occlusion.
  s->mesh_table[slot * kMeshStride + 3] = make_float4(1e30f, 0.0f, static_cast<float>(L), static_cast<float>(m.level_base));
  s->mesh_table[slot * kMeshStride + 5] = make_float4(1, 1, 0, 0);
  s->mesh_slots = std::max(s->mesh_slots, slot + 1);
  s->dirty_meshlets = s->dirty_table = true;
  *slot_out = slot;
  return VG_OK;
}

namespace {

// Instances of the used meshes other than `skip`.
size_t live_instances(const vg_scene* s, uint32_t skip) {
  size_t live = 0;
  for (uint32_t k = 0; k < s->slots; k++)
    if (s->meshes[k].used && k != skip) live += s->meshes[k].instance_count;
  return live;
}

// Moves every used mesh's instance range (but `skip`'s, which is dropped) to the front of the columns, in slot
// order, and rebases its cells' first instances (exact: instance indices are below 2^24).
void repack_instances(vg_scene* s, uint32_t skip) {
  std::vector<float4> col[4];
  std::vector<float> islot;
  const size_t live = live_instances(s, skip);
  for (int k = 0; k < 4; k++) col[k].reserve(live);
  islot.reserve(live);
  for (uint32_t slot = 0; slot < s->slots; slot++) {
    MeshHost& m = s->meshes[slot];
    if (!m.used || slot == skip) continue;
    const uint32_t base = static_cast<uint32_t>(islot.size());
    for (int k = 0; k < 4; k++) col[k].insert(col[k].end(), s->i_col[k].begin() + m.instance_base, s->i_col[k].begin() + m.instance_base + m.instance_count);
    islot.insert(islot.end(), m.instance_count, static_cast<float>(slot));
    for (float4& rec : m.cell_rec) rec.z = static_cast<float>(static_cast<uint32_t>(rec.z) - m.instance_base + base);
    m.instance_base = base;
  }
  for (int k = 0; k < 4; k++) s->i_col[k].swap(col[k]);
  s->i_slot.swap(islot);
  if (skip < s->slots) {
    s->meshes[skip].instance_base = 0;
    s->meshes[skip].instance_count = 0;
    s->meshes[skip].cell_sphere.clear();
    s->meshes[skip].cell_rec.clear();
    s->meshes[skip].cell_bound.clear();
  }
  s->dirty_instances = s->dirty_cells = true;
}

}  // namespace

int vg_mesh_remove(vg_scene* s, uint32_t slot) {
  if (!s || slot >= s->slots || !s->meshes[slot].used) return VG_E_INVALID;
  std::fill(s->phase1_ready.begin(), s->phase1_ready.end(), 0);  // instance slots may move
  MeshHost& m = s->meshes[slot];
  const size_t live = live_instances(s, slot);
  if (s->i_col[0].size() > 2 * live) {
    repack_instances(s, slot);
  } else {  // its range stays dead until a later repack
    m.instance_count = 0;
    m.cell_sphere.clear();
    m.cell_rec.clear();
    m.cell_bound.clear();
  }
  m.used = false;
  s->mesh_table[slot * kMeshStride + 5] = make_float4(0, 0, 0, 0);  // writeFlags({enabled: 0})
  s->dirty_table = s->dirty_cells = true;
  return VG_OK;
}

int vg_instances_set(vg_scene* s, uint32_t slot, const float* matrices, const float* colors, uint32_t count,
                     uint32_t* gpu_to_input) {
  (void)colors;  // tints are raster-side data (GeometryPool.ts:490-491); the cut does not read them
  if (!s || slot >= s->slots || !s->meshes[slot].used || (!matrices && count)) return VG_E_INVALID;
  std::fill(s->phase1_ready.begin(), s->phase1_ready.end(), 0);  // instance slots move: recorded rejects no longer apply
  vg_ctx* ctx = s->ctx;
  MeshHost& m = s->meshes[slot];
  if (count != m.instance_count || m.instance_count == 0) {
    // The mesh's old range dies; repack first when the dead part would pass half of the columns.
    const size_t live = live_instances(s, slot);
    if (s->i_col[0].size() + count > 2 * (live + count) || s->i_col[0].size() + count > kMaxInstancesPerScene) repack_instances(s, slot);
    const size_t total = s->i_col[0].size();
    if (total + count > kMaxInstancesPerScene) return fail(ctx, VG_E_CAPACITY, "vg_instances_set: more than 2^24 instances in a scene");
    m.instance_base = static_cast<uint32_t>(total);
    for (int k = 0; k < 4; k++) s->i_col[k].resize(total + count);
    s->i_slot.resize(total + count, static_cast<float>(slot));
  }
  m.instance_count = count;
  std::vector<uint32_t> order = morton_order(matrices, count, m.bounding);
  std::vector<float> gpu(static_cast<size_t>(count) * 16);
  for (uint32_t g = 0; g < count; g++) {
    const float* src = matrices + static_cast<size_t>(order[g]) * 16;
    std::memcpy(&gpu[static_cast<size_t>(g) * 16], src, sizeof(float) * 16);
    for (int k = 0; k < 4; k++) s->i_col[k][m.instance_base + g] = make_float4(src[k * 4], src[k * 4 + 1], src[k * 4 + 2], src[k * 4 + 3]);
  }
  if (gpu_to_input) std::memcpy(gpu_to_input, order.data(), sizeof(uint32_t) * count);
  const uint32_t cells = (count + kCellSize - 1) / kCellSize;
  m.cell_sphere.resize(cells);
  m.cell_rec.resize(cells);
  m.cell_bound.resize(cells);
  for (uint32_t c = 0; c < cells; c++) {
    const uint32_t first = c * kCellSize, end = std::min(count, (c + 1) * kCellSize);
    compute_cell(gpu, first, end, m.bounding, slot, m.instance_base, m.cell_sphere[c], m.cell_rec[c]);
    m.cell_bound[c] = cell_bound(gpu, first, end, m.bounding, m.cull_bounds, m.cell_sphere[c]);
  }
  s->dirty_instances = s->dirty_cells = true;
  return VG_OK;
}

int vg_mesh_flags(vg_scene* s, uint32_t slot, float max_draw_distance, float min_pixel_radius, uint32_t enabled, uint32_t shadow,
                  uint32_t occlusion) {
  if (!s || slot >= s->slots || !s->meshes[slot].used) return VG_E_INVALID;
  float4& levers = s->mesh_table[slot * kMeshStride + 3];
  levers.x = max_draw_distance;
  levers.y = min_pixel_radius;
  s->mesh_table[slot * kMeshStride + 5] = make_float4(enabled ? 1.0f : 0.0f, shadow ? 1.0f : 0.0f, occlusion ? 1.0f : 0.0f, 0.0f);
  s->dirty_table = true;
  return VG_OK;
}

int vg_hzb_set(vg_scene* s, uint32_t view_slot, const float* depth, uint32_t width, uint32_t height) {
  if (!s || !depth || view_slot >= s->ctx->cfg.max_views) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "vg_hzb_set: host-only context");
  const uint32_t max_size = ctx->cfg.hzb_max_size;
  if (width == 0 || height == 0 || width > max_size || height > max_size)
    return fail(ctx, VG_E_INVALID, "vg_hzb_set: depth size must be within [1, hzb_max_size]");
  // Layout: header + one region per level sized for max_size (OcclusionCulling.ts:92-113, 284-300).
  std::vector<uint32_t> offsets;
  uint32_t total = kHzbHeader;
  for (uint32_t dim = max_size;;) {
    offsets.push_back(total);
    total += dim * dim;
    if (dim == 1) break;
    dim = (dim + 1) / 2;
  }
  std::vector<float> header(kHzbHeader, 0.0f);
  uint32_t lw = width, lh = height, levels = 0;
  std::vector<uint32_t> ws, hs;
  for (size_t i = 0; i < offsets.size(); i++) {
    header[i * 4] = static_cast<float>(offsets[i]);
    header[i * 4 + 1] = static_cast<float>(lw);
    header[i * 4 + 2] = static_cast<float>(lh);
    ws.push_back(lw);
    hs.push_back(lh);
    levels++;
    if (lw == 1 && lh == 1) break;
    lw = lw > 1 ? (lw + 1) / 2 : 1;
    lh = lh > 1 ? (lh + 1) / 2 : 1;
  }
  header[kHzbLevelCountSlot] = static_cast<float>(levels);
  if (!s->d_hzbs[view_slot].ptr) {
    if (int rc = ensure(s, s->d_hzbs[view_slot], sizeof(float) * total)) return rc;
    s->h_hzb_ptrs[view_slot] = static_cast<float*>(s->d_hzbs[view_slot].ptr);
    VG_HIP(ctx, hipMemcpyAsync(static_cast<float**>(s->d_hzb_ptrs.ptr) + view_slot, &s->h_hzb_ptrs[view_slot], sizeof(float*),
                               hipMemcpyHostToDevice, ctx->stream));
  }
  float* hzb = static_cast<float*>(s->d_hzbs[view_slot].ptr);
  hipLaunchKernelGGL(k_zero_words, dim3(std::min<uint32_t>((total - kHzbHeader + 255) / 256, 4096u)), dim3(256), 0, ctx->stream,
                     reinterpret_cast<uint32_t*>(hzb + kHzbHeader), total - kHzbHeader);
  VG_HIP(ctx, hipMemcpyAsync(hzb, header.data(), sizeof(float) * kHzbHeader, hipMemcpyHostToDevice, ctx->stream));
  VG_HIP(ctx, hipMemcpyAsync(hzb + offsets[0], depth, sizeof(float) * width * height, hipMemcpyHostToDevice, ctx->stream));
  for (uint32_t base = 0; base + 1 < levels; base += 5) {
    dim3 grid((ws[base] + 31) / 32, (hs[base] + 31) / 32);
    hipLaunchKernelGGL(k_hzb_pass, grid, dim3(256), 0, ctx->stream, hzb, base);
  }
  VG_HIP(ctx, hipGetLastError());
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->hzb_valid[view_slot] = 1;
  return VG_OK;
}

int vg_hzb_from_cuts(vg_scene* s, const vg_raster* rasters, uint32_t count) {
  if (!s || (!rasters && count)) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "vg_hzb_from_cuts: host-only context");
  if (count == 0) return VG_OK;
  const uint32_t max_size = ctx->cfg.hzb_max_size;
  if (count > kMaxRasters) return fail(ctx, VG_E_INVALID, "vg_hzb_from_cuts: more than 256 rasters in one call");
  std::vector<uint8_t> seen(ctx->cfg.max_views, 0);
  for (uint32_t i = 0; i < count; i++) {
    const vg_raster& q = rasters[i];
    if (q.view_slot >= ctx->cfg.max_views) return VG_E_INVALID;
    if (q.width == 0 || q.height == 0 || q.width > max_size || q.height > max_size)
      return fail(ctx, VG_E_INVALID, "vg_hzb_from_cuts: depth size must be within [1, hzb_max_size]");
    if (!s->d_draws[q.view_slot].ptr) return fail(ctx, VG_E_INVALID, "vg_hzb_from_cuts: the slot has no cut yet");
    if (seen[q.view_slot]++) return fail(ctx, VG_E_INVALID, "vg_hzb_from_cuts: a view slot appears twice");
  }
  if (int rc = upload_scene(s)) return rc;
  uint32_t total = kHzbHeader;
  for (uint32_t dim = max_size;; dim = (dim + 1) / 2) {
    total += dim * dim;
    if (dim == 1) break;
  }
  // The rasters' records: camera, draw list, entry count (the cut's, or with its second phase's), HZB.
  std::vector<RasterView> views(count);
  uint32_t max_w = 0, max_h = 0;
  for (uint32_t i = 0; i < count; i++) {
    const vg_raster& q = rasters[i];
    const uint32_t slot = q.view_slot;
    if (!s->d_hzbs[slot].ptr) {
      if (int rc = ensure(s, s->d_hzbs[slot], sizeof(float) * total)) return rc;
      s->h_hzb_ptrs[slot] = static_cast<float*>(s->d_hzbs[slot].ptr);
      VG_HIP(ctx, hipMemcpyAsync(static_cast<float**>(s->d_hzb_ptrs.ptr) + slot, &s->h_hzb_ptrs[slot], sizeof(float*),
                                 hipMemcpyHostToDevice, ctx->stream));
    }
    RasterView& v = views[i];
    std::memset(&v, 0, sizeof(v));
    std::memcpy(v.vp, q.view_proj, sizeof(v.vp));
    v.draws = static_cast<const uint4*>(s->d_draws[slot].ptr);
    // Camera entries: [0, drawn) of the cut, or [0, min(selected, max)) when its second phase appended more.
    const bool both = q.with_phase2 && s->phase2_ran[slot] && s->d_counters2.ptr;
    const uint32_t* counters = static_cast<const uint32_t*>(both ? s->d_counters2.ptr : s->d_counters.ptr) +
                               static_cast<size_t>(slot) * s->counter_stride;
    v.count = counters + 2 + (both ? kSelected : kDrawn);
    v.hzb = static_cast<float*>(s->d_hzbs[slot].ptr);
    v.width = q.width;
    v.height = q.height;
    max_w = std::max(max_w, q.width);
    max_h = std::max(max_h, q.height);
  }
  if (int rc = ensure(s, s->d_raster_views, sizeof(RasterView) * kMaxRasters)) return rc;
  // From pageable memory: the copy has taken the records when it returns.
  VG_HIP(ctx, hipMemcpyAsync(s->d_raster_views.ptr, views.data(), sizeof(RasterView) * count, hipMemcpyHostToDevice, ctx->stream));
  RasterParams r;
  std::memset(&r, 0, sizeof(r));
  for (int k = 0; k < 4; k++) r.col[k] = static_cast<const float4*>(s->d_col[k].ptr);
  r.first = static_cast<const uint32_t*>(s->d_first.ptr);
  r.tris = static_cast<const uint32_t*>(s->d_tris.ptr);
  r.vlist = static_cast<const uint2*>(s->d_vlist.ptr);
  r.tri = static_cast<const uint32_t*>(s->d_tri.ptr);
  r.vpos = static_cast<const float4*>(s->d_vpos.ptr);
  r.views = static_cast<const RasterView*>(s->d_raster_views.ptr);
  r.rasters = count;
  r.clamp = ctx->cfg.max_meshlets;
  r.max_size = max_size;
  r.clear_texels = total - kHzbHeader;
  constexpr uint32_t kChunkCapacity = 1u << 20;  // 32 MB of queued rectangles; beyond it the waves fill them
  if (!s->d_raster_chunks.ptr) {
    if (int rc = ensure(s, s->d_raster_chunks, sizeof(RasterChunk) * kChunkCapacity)) return rc;
    if (int rc = ensure(s, s->d_raster_count, sizeof(uint32_t))) return rc;
  }
  r.chunks = static_cast<RasterChunk*>(s->d_raster_chunks.ptr);
  r.chunk_count = static_cast<uint32_t*>(s->d_raster_count.ptr);
  r.chunk_capacity = kChunkCapacity;
  hipLaunchKernelGGL(k_hzb_header, dim3(count), dim3(64), 0, ctx->stream, r);
  hipLaunchKernelGGL(k_hzb_clear, dim3(std::min<uint32_t>((total - kHzbHeader + 255) / 256, std::max<uint32_t>(4096u / count, 64u)), count),
                     dim3(256), 0, ctx->stream, r);
  if (!s->m_verts.empty()) {
    const dim3 grid(static_cast<uint32_t>(ctx->cus) * kRasterBlocksPerCu);
    std::vector<hipEvent_t> ev;
    if (s->profile) {
      ev.assign(kKernelSlots + 1, nullptr);
      ev[kRasterSlot] = take_event(s);
      ev[kRasterSlot + 1] = take_event(s);
      if (ev[kRasterSlot]) (void)hipEventRecord(ev[kRasterSlot], ctx->stream);
    }
    hipLaunchKernelGGL(k_raster<kRasterSub>, grid, dim3(256), 0, ctx->stream, r);
    hipLaunchKernelGGL(k_raster_chunks, grid, dim3(256), 0, ctx->stream, r);
    if (s->profile) {
      if (ev[kRasterSlot + 1]) (void)hipEventRecord(ev[kRasterSlot + 1], ctx->stream);
      s->pending.push_back(ev);
    }
  }
  // The pyramid, 5 levels per pass, every raster in one launch per pass (grids sized for the largest image).
  uint32_t lw = max_w, lh = max_h, levels = 0;
  std::vector<uint32_t> ws, hs;
  for (uint32_t dim = max_size;; dim = (dim + 1) / 2) {
    ws.push_back(lw);
    hs.push_back(lh);
    levels++;
    if ((lw == 1 && lh == 1) || dim == 1) break;
    lw = lw > 1 ? (lw + 1) / 2 : 1;
    lh = lh > 1 ? (lh + 1) / 2 : 1;
  }
  for (uint32_t base = 0; base + 1 < levels; base += 5) {
    dim3 grid((ws[base] + 31) / 32, (hs[base] + 31) / 32, count);
    hipLaunchKernelGGL(k_hzb_pass_rasters, grid, dim3(256), 0, ctx->stream, r, base);
  }
  VG_HIP(ctx, hipGetLastError());
  for (uint32_t i = 0; i < count; i++) s->hzb_valid[rasters[i].view_slot] = 1;
  return VG_OK;
}

int vg_hzb_from_cut(vg_scene* s, uint32_t view_slot, const float* view_proj, uint32_t width, uint32_t height, uint32_t with_phase2) {
  if (!s || !view_proj) return VG_E_INVALID;
  vg_raster q;
  q.view_slot = view_slot;
  q.width = width;
  q.height = height;
  q.with_phase2 = with_phase2;
  std::memcpy(q.view_proj, view_proj, sizeof(q.view_proj));
  return vg_hzb_from_cuts(s, &q, 1);
}

int vg_hzb_read(vg_scene* s, uint32_t view_slot, float* dst, uint64_t capacity_floats, uint64_t* size_floats) {
  if (!s || view_slot >= s->ctx->cfg.max_views) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  const uint64_t n = s->d_hzbs[view_slot].ptr ? s->d_hzbs[view_slot].bytes / sizeof(float) : 0;
  if (size_floats) *size_floats = n;
  if (dst && n) {
    if (capacity_floats < n) return VG_E_CAPACITY;
    VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    VG_HIP(ctx, hipMemcpy(dst, s->d_hzbs[view_slot].ptr, sizeof(float) * n, hipMemcpyDeviceToHost));
  }
  return VG_OK;
}

int vg_hzb_invalidate(vg_scene* s, uint32_t view_slot) {
  if (!s || view_slot >= s->ctx->cfg.max_views) return VG_E_INVALID;
  s->hzb_valid[view_slot] = 0;
  return VG_OK;
}

// After a cut: a view that asked for two-phase occlusion may take its second phase next.
static void mark_phase1(vg_scene* s, const vg_view* views, uint32_t count) {
  for (uint32_t i = 0; i < count; i++) {
    s->phase1_ready[views[i].slot] = views[i].occlusion == 2 ? 1 : 0;
    s->phase2_ran[views[i].slot] = 0;
  }
}

int vg_cut(vg_scene* s, const vg_view* views, uint32_t count) {
  if (!s) return VG_E_INVALID;
  if (int rc = prepare_cut(s, &s->lane, views, count, false)) return rc;
  if (int rc = enqueue_cut(s, &s->lane, count, false)) return rc;
  mark_phase1(s, views, count);
  return VG_OK;
}

int vg_cut_graph(vg_scene* s, const vg_view* views, uint32_t count) {
  if (!s) return VG_E_INVALID;
  if (int rc = prepare_cut(s, &s->lane, views, count, false)) return rc;
  if (int rc = launch_graph(s, &s->lane, count)) return rc;
  mark_phase1(s, views, count);
  return VG_OK;
}

int vg_cut_static(vg_scene* s, const vg_view* views, uint32_t count) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (count > kStaticViews) return fail(ctx, VG_E_CAPACITY, "vg_cut_static: more than 64 views (the size its static scratch is laid out for)");
  if (int rc = check_views(ctx, views, count, "vg_cut_static")) return rc;
  if (!ctx->static_scratch) return fail(ctx, VG_E_DEVICE, "vg_cut_static: no static scratch on this context's device");
  if (ctx->cfg.expand)
    return fail(ctx, VG_E_INVALID, "vg_cut_static: contexts with the expand pass are not supported (its per-entry records do not fit the static scratch)");
  if (s->dirty_table || s->dirty_meshlets || s->dirty_instances || s->dirty_cells)
    return fail(ctx, VG_E_INVALID, "vg_cut_static: the scene has changes that are not on the device yet (a vg_cut uploads them)");
  const uint32_t reject_words = static_cast<uint32_t>((s->i_col[0].size() + 31) / 32 + 1);
  for (uint32_t i = 0; i < count; i++) {
    if (!s->d_draws[views[i].slot].ptr)
      return fail(ctx, VG_E_INVALID, "vg_cut_static: a view slot without outputs yet (a vg_cut of the slot allocates them)");
    if (views[i].occlusion == 2 && (!s->d_phase.ptr || s->reject_words < reject_words))
      return fail(ctx, VG_E_INVALID, "vg_cut_static: two-phase occlusion needs the scene's two-phase buffers (a vg_cut with occlusion 2 allocates them)");
  }
  // Layout of the scratch for this call, 256-byte aligned pieces; the work list takes what is left.
  const uint32_t bins = ctx->cfg.budget_bins;
  const size_t cells = std::max<size_t>(s->c_sphere.size(), 1);
  size_t used = 0;
  auto take = [&](size_t bytes) {
    const size_t at = used;
    used += (bytes + 255) & ~static_cast<size_t>(255);
    return at;
  };
  const size_t at_views = take(sizeof(ViewDev) * count), at_queue_count = take(64);
  const size_t at_work_counter = take(sizeof(unsigned long long) * kWorkParts * kWorkCounterSpacing), at_flags = take(64);
  const size_t at_thresholds = take(sizeof(float) * count * bins), at_hist = take(sizeof(unsigned long long) * count * (bins + 1));
  const size_t at_plan = take(sizeof(unsigned long long) * (2 * static_cast<size_t>(count) + 1));
  const size_t at_queue = take(sizeof(uint2) * cells * count);
  const size_t per_item = sizeof(uint4) + sizeof(unsigned long long) + sizeof(float);
  const size_t left = used + 3 * 256 < kStaticScratchBytes ? kStaticScratchBytes - used - 3 * 256 : 0;
  const size_t per_part = std::min<size_t>(left / per_item / kWorkParts, kMaxWorkItems / kWorkParts);
  if (per_part == 0)
    return fail(ctx, VG_E_CAPACITY, "vg_cut_static: " + std::to_string(count) + " views x " + std::to_string(cells) +
                                        " cells do not fit its 64 MiB static scratch");
  const size_t items = per_part * kWorkParts;
  const size_t at_work = take(sizeof(uint4) * items), at_start = take(sizeof(unsigned long long) * items), at_scale = take(sizeof(float) * items);
  if (used > kStaticScratchBytes) return fail(ctx, VG_E_CAPACITY, "vg_cut_static: the launch does not fit its static scratch");
  char* base = static_cast<char*>(ctx->static_scratch);
  Lane lane;
  lane.stream = ctx->stream;
  lane.borrowed = true;
  auto piece = [&](DevBuf& b, size_t at, size_t bytes) {
    b.ptr = base + at;
    b.bytes = bytes;
  };
  piece(lane.d_views, at_views, sizeof(ViewDev) * count);
  piece(lane.d_cell_queue_count, at_queue_count, 64);
  piece(lane.d_work_counter, at_work_counter, sizeof(unsigned long long) * kWorkParts * kWorkCounterSpacing);
  piece(lane.d_flags, at_flags, 64);
  piece(lane.d_budget_thresholds, at_thresholds, sizeof(float) * count * bins);
  piece(lane.d_budget_hist, at_hist, sizeof(unsigned long long) * count * (bins + 1));
  piece(lane.d_expand_plan, at_plan, sizeof(unsigned long long) * (2 * static_cast<size_t>(count) + 1));
  piece(lane.d_cell_queue, at_queue, sizeof(uint2) * cells * count);
  piece(lane.d_work, at_work, sizeof(uint4) * items);
  piece(lane.d_work_start, at_start, sizeof(unsigned long long) * items);
  piece(lane.d_work_scale, at_scale, sizeof(float) * items);
  lane.cell_queue_capacity = cells * count;
  lane.work_capacity = static_cast<uint32_t>(items);
  lane.view_capacity = count;
  // Everything the kernels need zeroed (queue count, work counters, flags, histograms) k_begin clears.
  lane.launch_budget = stage_views(s, views, count, false, reinterpret_cast<ViewDev*>(ctx->static_views.data()), ctx->static_thresholds.data());
  if (lane.launch_budget)
    VG_HIP(ctx, hipMemcpyAsync(lane.d_budget_thresholds.ptr, ctx->static_thresholds.data(), sizeof(float) * count * bins,
                               hipMemcpyHostToDevice, ctx->stream));
  VG_HIP(ctx, hipMemcpyAsync(lane.d_views.ptr, ctx->static_views.data(), sizeof(ViewDev) * count, hipMemcpyHostToDevice, ctx->stream));
  if (int rc = enqueue_cut(s, &lane, count, false)) return rc;
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  uint32_t flags[4] = {0, 0, 0, 0};
  VG_HIP(ctx, hipMemcpy(flags, lane.d_flags.ptr, sizeof(flags), hipMemcpyDeviceToHost));
  mark_phase1(s, views, count);
  if (flags[3])
    return fail(ctx, VG_E_CAPACITY, "vg_cut_static: the work list (" + std::to_string(items) +
                                        " items in the static scratch) overflowed; results are incomplete");
  if (flags[2]) return fail(ctx, VG_E_DEVICE, "vg_cut_static: internal error in the meshlet pass");
  return VG_OK;
}

int vg_cut_phase2(vg_scene* s, const vg_view* views, uint32_t count) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!views || count == 0) return fail(ctx, VG_E_INVALID, "vg_cut_phase2: no views");
  for (uint32_t i = 0; i < count; i++)
    if (views[i].slot >= ctx->cfg.max_views || views[i].occlusion != 2 || !s->phase1_ready[views[i].slot])
      return fail(ctx, VG_E_INVALID, "vg_cut_phase2: every view needs occlusion = 2 and its slot's cut (vg_cut) just before");
  if (int rc = prepare_cut(s, &s->lane, views, count, false)) return rc;
  if (int rc = enqueue_phase2(s, &s->lane, count)) return rc;
  for (uint32_t i = 0; i < count; i++) {
    s->phase1_ready[views[i].slot] = 0;
    s->phase2_ran[views[i].slot] = 1;
  }
  return VG_OK;
}

int vg_read_phase2(vg_scene* s, uint32_t view_slot, uint32_t* counters, uint32_t* draw_args) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  if (view_slot >= ctx->cfg.max_views) return fail(ctx, VG_E_INVALID, "vg_read_phase2: view slot out of range");
  if (!s->d_phase.ptr) return fail(ctx, VG_E_INVALID, "vg_read_phase2: no view of this scene asked for two-phase occlusion");
  if (counters)
    VG_HIP(ctx, hipMemcpyAsync(counters, static_cast<uint32_t*>(s->d_counters2.ptr) + static_cast<size_t>(view_slot) * s->counter_stride,
                               sizeof(uint32_t) * s->counter_stride, hipMemcpyDeviceToHost, ctx->stream));
  if (draw_args)
    VG_HIP(ctx, hipMemcpyAsync(draw_args, static_cast<uint32_t*>(s->d_args2.ptr) + static_cast<size_t>(view_slot) * s->args_stride,
                               sizeof(uint32_t) * s->args_stride, hipMemcpyDeviceToHost, ctx->stream));
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return VG_OK;
}

int vg_phase_rejects(vg_scene* s, uint32_t view_slot, uint32_t* meshlets, uint32_t* dropped) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  if (view_slot >= ctx->cfg.max_views || !s->d_phase.ptr) return fail(ctx, VG_E_INVALID, "vg_phase_rejects: no two-phase data for this slot");
  uint32_t w[kPhaseWords];
  VG_HIP(ctx, hipMemcpyAsync(w, static_cast<uint32_t*>(s->d_phase.ptr) + static_cast<size_t>(view_slot) * kPhaseWords, sizeof(w),
                             hipMemcpyDeviceToHost, ctx->stream));
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (meshlets) *meshlets = w[P_REJECTS];
  if (dropped) *dropped = w[P_OVERFLOW];
  return VG_OK;
}

int vg_sync(vg_scene* s) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (bool& p : s->lane.ring_pending) p = false;
  uint32_t flags[4] = {0, 0, 0, 0};
  VG_HIP(ctx, hipMemcpy(flags, s->lane.d_flags.ptr, sizeof(flags), hipMemcpyDeviceToHost));
  if (flags[3]) return fail(ctx, VG_E_CAPACITY, "vg_cut: work list capacity exceeded; results are incomplete");
  if (flags[2]) return fail(ctx, VG_E_DEVICE, "vg_cut: internal error in the meshlet pass");
  return VG_OK;
}

int vg_read(vg_scene* s, uint32_t view_slot, uint32_t* counters, uint32_t* draw_args, uint32_t* draw_list, uint32_t max_entries,
            vg_metrics* metrics) {
  if (!s || view_slot >= s->ctx->cfg.max_views) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  return read_slot(s, ctx->stream, view_slot, counters, draw_args, draw_list, max_entries, metrics);
}

int vg_index_buffer(vg_scene* s, uint32_t view_slot, uint32_t cut, void** device_ptr, uint64_t* capacity) {
  if (!s || cut > 1) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  if (!ctx->cfg.expand) return fail(ctx, VG_E_INVALID, "vg_index_buffer: the context was created without the expand pass");
  if (view_slot >= ctx->cfg.max_views) return fail(ctx, VG_E_INVALID, "vg_index_buffer: view slot out of range");
  if (int rc = ensure_slot(s, view_slot)) return rc;
  const DevBuf& b = s->d_indices[2 * view_slot + cut];
  if (device_ptr) *device_ptr = b.ptr;
  if (capacity) *capacity = static_cast<uint64_t>(cut ? ctx->cfg.max_shadow_triangles : ctx->cfg.max_triangles) * 3;
  return VG_OK;
}

int vg_index_read(vg_scene* s, uint32_t view_slot, uint32_t cut, uint64_t first, uint64_t count, uint32_t* dst) {
  void* ptr = nullptr;
  uint64_t capacity = 0;
  if (int rc = vg_index_buffer(s, view_slot, cut, &ptr, &capacity)) return rc;
  if (first > capacity || count > capacity - first) return fail(s->ctx, VG_E_INVALID, "vg_index_read: range outside the index buffer");
  if (count == 0) return VG_OK;
  if (!dst) return VG_E_INVALID;
  VG_HIP(s->ctx, hipMemcpyAsync(dst, static_cast<uint32_t*>(ptr) + first, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s->ctx->stream));
  VG_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  return VG_OK;
}

int vg_state_get(vg_scene* s, uint32_t view_slot, uint32_t* state) {
  if (!s || !state || view_slot >= s->ctx->cfg.max_views) return VG_E_INVALID;
  if (!s->ctx->stream) return fail(s->ctx, VG_E_DEVICE, "host-only context");
  VG_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  VG_HIP(s->ctx, hipMemcpy(state, static_cast<uint32_t*>(s->d_counters.ptr) + static_cast<size_t>(view_slot) * s->counter_stride +
                                      state_base(s->slots),
                           sizeof(uint32_t) * kStateWords, hipMemcpyDeviceToHost));
  return VG_OK;
}

int vg_state_set(vg_scene* s, uint32_t view_slot, const uint32_t* state) {
  if (!s || !state || view_slot >= s->ctx->cfg.max_views) return VG_E_INVALID;
  if (!s->ctx->stream) return fail(s->ctx, VG_E_DEVICE, "host-only context");
  VG_HIP(s->ctx, hipStreamSynchronize(s->ctx->stream));
  VG_HIP(s->ctx, hipMemcpy(static_cast<uint32_t*>(s->d_counters.ptr) + static_cast<size_t>(view_slot) * s->counter_stride +
                               state_base(s->slots),
                           state, sizeof(uint32_t) * kStateWords, hipMemcpyHostToDevice));
  return VG_OK;
}

int vg_scene_pool_read(const vg_scene* s, uint32_t buffer, void* dst, uint64_t capacity_bytes, uint64_t* size_bytes) {
  if (!s || buffer > VG_POOL_MESHLET_VERTICES) return VG_E_INVALID;
  std::vector<float> f;
  const void* src = nullptr;
  uint64_t bytes = 0;
  switch (buffer) {
    case VG_POOL_MESH_TABLE:
      src = s->mesh_table.data();
      bytes = sizeof(float4) * s->mesh_table.size();
      break;
    case VG_POOL_INSTANCES: {
      const size_t n = s->i_col[0].size();
      f.resize(n * 20);
      for (size_t i = 0; i < n; i++) {
        for (int k = 0; k < 4; k++) std::memcpy(&f[i * 20 + k * 4], &s->i_col[k][i], sizeof(float4));
        f[i * 20 + 16] = f[i * 20 + 17] = f[i * 20 + 18] = 1.0f;
        f[i * 20 + 19] = s->i_slot[i];
      }
      src = f.data();
      bytes = sizeof(float) * f.size();
      break;
    }
    case VG_POOL_CELLS: {
      f.clear();
      for (const MeshHost& m : s->meshes) {
        if (!m.used) continue;
        for (size_t c = 0; c < m.cell_sphere.size(); c++) {
          const float4 a = m.cell_sphere[c], b = m.cell_rec[c];
          const float v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
          f.insert(f.end(), v, v + 8);
        }
      }
      src = f.data();
      bytes = sizeof(float) * f.size();
      break;
    }
    case VG_POOL_MESHLET_BOUNDS: {
      const size_t n = s->m_cull.size();
      f.resize(n * 16);
      for (size_t i = 0; i < n; i++) {
        std::memcpy(&f[i * 16], &s->m_lod[i], sizeof(float4));
        std::memcpy(&f[i * 16 + 4], &s->m_parent[i], sizeof(float4));
        f[i * 16 + 8] = s->m_err[i].x;
        f[i * 16 + 9] = s->m_err[i].y;
        f[i * 16 + 10] = f[i * 16 + 11] = 0.0f;
        std::memcpy(&f[i * 16 + 12], &s->m_cull[i], sizeof(float4));
      }
      src = f.data();
      bytes = sizeof(float) * f.size();
      break;
    }
    case VG_POOL_MESHLET_INFO:
      src = s->m_info.data();
      bytes = sizeof(uint32_t) * s->m_info.size();
      break;
    case VG_POOL_VERTICES:
      src = s->m_verts.data();
      bytes = sizeof(float4) * s->m_verts.size();
      break;
    case VG_POOL_MESHLET_VERTICES:
      src = s->m_mverts.data();
      bytes = sizeof(uint32_t) * s->m_mverts.size();
      break;
    default:
      src = s->m_tri.data();
      bytes = sizeof(uint32_t) * s->m_tri.size();
      break;
  }
  if (size_bytes) *size_bytes = bytes;
  if (dst) {
    if (capacity_bytes < bytes) return VG_E_CAPACITY;
    if (bytes) std::memcpy(dst, src, bytes);
  }
  return VG_OK;
}

uint32_t vg_scene_mesh_slots(const vg_scene* s) { return s ? s->mesh_slots : 0; }

int vg_budget_thresholds(float lo, float hi, uint32_t bins, float* out) {
  if (!out || bins == 0 || !(lo > 0) || !(hi >= lo)) return VG_E_INVALID;
  for (uint32_t b = 0; b < bins; b++)
    out[b] = bins == 1 ? lo : static_cast<float>(static_cast<double>(lo) * std::pow(static_cast<double>(hi) / lo, static_cast<double>(b) / (bins - 1)));
  out[bins - 1] = hi;
  return VG_OK;
}

int vg_profile(vg_scene* s, uint32_t enable) {
  if (!s) return VG_E_INVALID;
  s->profile = enable != 0;
  return VG_OK;
}

int vg_profile_read(vg_scene* s, double* kernel_ms, uint32_t max_kernels, uint64_t* sequences, const char** names) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (!ctx->stream) return fail(ctx, VG_E_DEVICE, "host-only context");
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& seq : s->pending) {
    for (int i = 0; i < kKernelSlots; i++) {
      float ms = 0;
      if (seq[i] && seq[i + 1] && hipEventElapsedTime(&ms, seq[i], seq[i + 1]) == hipSuccess) s->kernel_ms[i] += ms;
    }
    for (hipEvent_t e : seq)
      if (e) s->event_pool.push_back(e);
    if (seq[0]) s->profiled_sequences++;  // launch sequences of cuts (raster records time only its own slot)
  }
  s->pending.clear();
  for (uint32_t i = 0; i < max_kernels && i < static_cast<uint32_t>(kKernelSlots); i++) {
    if (kernel_ms) kernel_ms[i] = s->kernel_ms[i];
    if (names) names[i] = kKernelNames[i];
  }
  if (sequences) *sequences = s->profiled_sequences;
  for (double& v : s->kernel_ms) v = 0;
  s->profiled_sequences = 0;
  return VG_OK;
}

int vg_footprint(vg_scene* s, const vg_view* views, uint32_t count, uint64_t* instances, uint64_t* meshlets) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (int rc = prepare_cut(s, &s->lane, views, count, false)) return rc;
  const size_t iw = (s->i_col[0].size() + 31) / 32 + 1, mw = (s->m_cull.size() + 31) / 32 + 1;
  int rc = 0;
  if ((rc = ensure(s, s->d_fp_instances, iw * 4)) || (rc = ensure(s, s->d_fp_meshlets, mw * 4)) || (rc = ensure(s, s->d_fp_counts, 16)))
    return rc;
  VG_HIP(ctx, hipMemsetAsync(s->d_fp_instances.ptr, 0, iw * 4, ctx->stream));
  VG_HIP(ctx, hipMemsetAsync(s->d_fp_meshlets.ptr, 0, mw * 4, ctx->stream));
  VG_HIP(ctx, hipMemsetAsync(s->d_fp_counts.ptr, 0, 16, ctx->stream));
  s->footprint = true;
  rc = enqueue_cut(s, &s->lane, count, false);
  s->footprint = false;
  if (rc) return rc;
  auto* counts = static_cast<unsigned long long*>(s->d_fp_counts.ptr);
  hipLaunchKernelGGL(k_popcount, dim3(256), dim3(256), 0, ctx->stream, static_cast<const uint32_t*>(s->d_fp_instances.ptr),
                     static_cast<uint32_t>(iw), counts);
  hipLaunchKernelGGL(k_popcount, dim3(256), dim3(256), 0, ctx->stream, static_cast<const uint32_t*>(s->d_fp_meshlets.ptr),
                     static_cast<uint32_t>(mw), counts + 1);
  VG_HIP(ctx, hipGetLastError());
  unsigned long long h[2];
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  VG_HIP(ctx, hipMemcpy(h, counts, sizeof(h), hipMemcpyDeviceToHost));
  if (instances) *instances = h[0];
  if (meshlets) *meshlets = h[1];
  return vg_sync(s);
}

int vg_mesh_instance_base(const vg_scene* s, uint32_t slot, uint32_t* base, uint32_t* count) {
  if (!s || slot >= s->slots || !s->meshes[slot].used) return VG_E_INVALID;
  if (base) *base = s->meshes[slot].instance_base;
  if (count) *count = s->meshes[slot].instance_count;
  return VG_OK;
}

int vg_schedule_get(const vg_scene* s, vg_schedule* schedule) {
  if (!s || !schedule) return VG_E_INVALID;
  *schedule = s->schedule;
  return VG_OK;
}

int vg_schedule_set(vg_scene* s, const vg_schedule* schedule) {
  if (!s || !schedule) return VG_E_INVALID;
  if (!schedule_valid(*schedule))
    return fail(s->ctx, VG_E_INVALID, "vg_schedule_set: blocks per CU out of range or no meshlet-pass variant for that occupancy");
  s->schedule = *schedule;
  return VG_OK;
}

// Times candidate schedules by coordinate descent: the meshlet-pass variant and grid first (instance
// grid fixed), then the instance grid. Each candidate runs one untimed cut, then `repeats` cuts timed
// one by one with events on the context's stream; the median counts. The slots' counters (controller
// state included) are restored before every candidate and at the end.
int vg_tune(vg_scene* s, const vg_view* views, uint32_t count, uint32_t repeats, vg_schedule* best, float* best_ms) {
  if (!s) return VG_E_INVALID;
  vg_ctx* ctx = s->ctx;
  if (repeats == 0 || repeats > 64) return fail(ctx, VG_E_INVALID, "vg_tune: repeats must be 1-64");
  if (int rc = prepare_cut(s, &s->lane, views, count, false)) return rc;
  VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<uint8_t> saved(s->d_counters.bytes);
  VG_HIP(ctx, hipMemcpy(saved.data(), s->d_counters.ptr, saved.size(), hipMemcpyDeviceToHost));
  auto restore = [&]() -> int {
    VG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    VG_HIP(ctx, hipMemcpy(s->d_counters.ptr, saved.data(), saved.size(), hipMemcpyHostToDevice));
    return VG_OK;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VG_HIP(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(ctx, VG_E_DEVICE, "vg_tune: hipEventCreate");
  }
  const bool profile = s->profile;
  s->profile = false;
  const vg_schedule start = s->schedule;
  int rc = VG_OK;
  auto time_it = [&](const vg_schedule& sc, float* ms) -> int {
    if (int r = restore()) return r;
    s->schedule = sc;
    std::vector<float> t;
    for (uint32_t i = 0; i <= repeats; i++) {
      if (int r = prepare_cut(s, &s->lane, views, count, false)) return r;
      if (hipEventRecord(e0, ctx->stream) != hipSuccess) return fail(ctx, VG_E_DEVICE, "vg_tune: hipEventRecord");
      if (int r = enqueue_cut(s, &s->lane, count, false)) return r;
      if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        return fail(ctx, VG_E_DEVICE, "vg_tune: event");
      float e = 0;
      if (hipEventElapsedTime(&e, e0, e1) != hipSuccess) return fail(ctx, VG_E_DEVICE, "vg_tune: hipEventElapsedTime");
      if (i) t.push_back(e);  // the first cut warms up
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return VG_OK;
  };
  vg_schedule chosen = start;
  float chosen_ms = 0;
  rc = time_it(chosen, &chosen_ms);
  // Coordinate descent, two sweeps: the meshlet pass's occupancy variant and grid, then the instance
  // pass's grid. The incumbent is timed again in each sweep, so one lucky timing does not stick.
  static const uint32_t kMeshlet[][2] = {{4, 4}, {4, 8}, {5, 5}, {5, 10}, {6, 6}, {6, 12}, {7, 7}, {7, 14}};
  static const uint32_t kInstance[] = {4, 6, 7, 8, 12};
  for (int sweep = 0; sweep < 2 && !rc; sweep++) {
    if (sweep) rc = time_it(chosen, &chosen_ms);
    for (const auto& m : kMeshlet) {
      if (rc) break;
      vg_schedule c = chosen;
      c.meshlet_waves_per_simd = m[0];
      c.meshlet_blocks_per_cu = m[1];
      if (schedule_key(c) == schedule_key(chosen)) continue;
      float ms = 0;
      rc = time_it(c, &ms);
      if (!rc && ms < chosen_ms) {
        chosen = c;
        chosen_ms = ms;
      }
    }
    for (uint32_t b : kInstance) {
      if (rc) break;
      vg_schedule c = chosen;
      c.instance_blocks_per_cu = b;
      if (schedule_key(c) == schedule_key(chosen)) continue;
      float ms = 0;
      rc = time_it(c, &ms);
      if (!rc && ms < chosen_ms) {
        chosen = c;
        chosen_ms = ms;
      }
    }
  }
  const int restored = restore();
  s->profile = profile;
  s->schedule = rc ? start : chosen;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (restored) return restored;
  if (best) *best = chosen;
  if (best_ms) *best_ms = chosen_ms;
  return VG_OK;
}

// Test hook: evaluates the arithmetic the decisions depend on for n input pairs (see k_arith_probe).
int vg_arith_probe(const float* a, const float* b, float* out, uint32_t n) {
  if (!a || !b || !out || n == 0) return VG_E_INVALID;
  float *da = nullptr, *db = nullptr, *dout = nullptr;
  if (hipMalloc(&da, sizeof(float) * n) != hipSuccess) return VG_E_NOMEM;
  if (hipMalloc(&db, sizeof(float) * n) != hipSuccess) {
    (void)hipFree(da);
    return VG_E_NOMEM;
  }
  if (hipMalloc(&dout, sizeof(float) * n * 5) != hipSuccess) {
    (void)hipFree(da);
    (void)hipFree(db);
    return VG_E_NOMEM;
  }
  int rc = VG_OK;
  if (hipMemcpy(da, a, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(db, b, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess)
    rc = VG_E_DEVICE;
  if (rc == VG_OK) {
    hipLaunchKernelGGL(k_arith_probe, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, dout, n);
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(out, dout, sizeof(float) * n * 5, hipMemcpyDeviceToHost) != hipSuccess)
      rc = VG_E_DEVICE;
  }
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dout);
  return rc;
}

// Test hook: wave-scan probe over `groups` groups of 64 values (see k_scan_probe); out holds 128 per group.
int vg_scan_probe(const uint32_t* in, uint32_t* out, uint32_t groups) {
  if (!in || !out || groups == 0) return VG_E_INVALID;
  uint32_t *din = nullptr, *dout = nullptr;
  if (hipMalloc(&din, sizeof(uint32_t) * 64 * groups) != hipSuccess) return VG_E_NOMEM;
  if (hipMalloc(&dout, sizeof(uint32_t) * 128 * groups) != hipSuccess) {
    (void)hipFree(din);
    return VG_E_NOMEM;
  }
  int rc = VG_OK;
  if (hipMemcpy(din, in, sizeof(uint32_t) * 64 * groups, hipMemcpyHostToDevice) != hipSuccess) rc = VG_E_DEVICE;
  if (rc == VG_OK) {
    hipLaunchKernelGGL(k_scan_probe, dim3(groups), dim3(64), 0, 0, din, dout, groups);
    if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(out, dout, sizeof(uint32_t) * 128 * groups, hipMemcpyDeviceToHost) != hipSuccess)
      rc = VG_E_DEVICE;
  }
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

}  // extern "C"
