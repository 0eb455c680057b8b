// Host-side structures shared by the runtime (vgcut.hip) and the request server (vgserve.hip): contexts,
// resident scenes (one GeometryPool each, GeometryPool.ts:213-512) and launch lanes. Internal to
// libvgcut.so; the public interface is include/vgcut.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/vgcut.h"

namespace vgi {

constexpr uint32_t kInstanceStride = 16;             // floats per caller matrix
constexpr uint32_t kMaxMeshletsPerScene = 1u << 24;  // ids travel as f32 in the level tables and cells
constexpr uint32_t kMaxInstancesPerScene = 1u << 24;
constexpr uint32_t kStateWordCount = 8;               // controller words per view slot

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct MeshHost {
  bool used = false;
  uint32_t meshlet_base = 0, meshlet_count = 0;
  uint32_t level_base = 0, level_count = 0;  // level_base in float4 units of the mesh table
  float bounding[4], lod_bounds[4], cull_bounds[4];
  uint32_t instance_base = 0, instance_count = 0;
  std::vector<float4> cell_sphere, cell_rec, cell_bound;
};

struct ViewRecord;  // the device view record (vg::ViewDev), opaque here

// A launch lane: one HIP stream and the scratch one launch sequence needs (view records, the (view,
// cell) queue, the work list, flags, budget tables) plus a pinned staging ring for view records and the
// last captured graph. A scene's own cuts run on its default lane (the context's stream); the server
// runs batches of requests on lanes of its own, so batches on different lanes overlap on the device.
struct Lane {
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  // Most persistent blocks per CU its launches use (0: the scene's schedule as is): a server lane launched
  // beside others leaves room on each CU for their kernels.
  uint32_t max_blocks_per_cu = 0;
  DevBuf d_views, d_cell_queue, d_cell_queue_count, d_work, d_work_start, d_work_scale, d_work_counter, d_flags;
  DevBuf d_budget_thresholds, d_budget_hist;
  DevBuf d_expand_plan;  // per (view, cut): first draw-list chunk of the expand pass; then the total
  DevBuf d_expand_records;  // per draw-list entry of every view slot's cuts: resolved destination (16 B)
  uint32_t view_capacity = 0;
  size_t cell_queue_capacity = 0;
  uint32_t work_capacity = 0;
  // Transient buffers (queue and work list) come from the caller (the server's memory pool) instead of
  // growing here; `borrowed` makes prepare_cut check their size instead of reallocating.
  bool borrowed = false;
  // Pinned staging ring for view records: a launch reuses a buffer only after its earlier copy ran, so
  // back-to-back launches do not wait for each other on the host.
  static constexpr int kRing = 16;
  void* h_ring[kRing] = {nullptr};
  float* h_ring_thresholds[kRing] = {nullptr};
  hipEvent_t ring_copied[kRing] = {nullptr};
  bool ring_pending[kRing] = {false};
  uint32_t ring_next = 0;
  bool launch_budget = false;
  uint64_t layout_version = 1;
  uint64_t device_bytes = 0;
  // captured launch sequence
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint32_t graph_views = 0;
  uint64_t graph_key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

}  // namespace vgi

struct vg_ctx {
  vg_config cfg;
  hipStream_t stream = nullptr;
  int cus = 256;
  std::string error;
  // vg_cut_static: the device address of the library's static scratch array on this context's device
  // (from hipGetSymbolAddress at vg_create), and host staging for its view records and budget tables,
  // sized at vg_create so that the call itself allocates nothing.
  void* static_scratch = nullptr;
  std::vector<unsigned char> static_views;
  std::vector<float> static_thresholds;
};

struct vg_scene {
  vg_ctx* ctx = nullptr;
  uint32_t slots = 64;
  std::vector<vgi::MeshHost> meshes;
  uint32_t mesh_slots = 0;  // highest used slot + 1
  uint32_t level_used = 0;  // float4 entries of level tables
  // host mirrors (device layout)
  std::vector<float4> mesh_table;
  std::vector<float4> m_cull, m_lod, m_parent;
  std::vector<float2> m_err;
  std::vector<uint32_t> m_tris;
  std::vector<uint32_t> m_info;  // host mirror of meshletInfo, firstIndex rebased (GeometryPool.ts:417-425)
  std::vector<uint32_t> m_first; // meshletInfo.y per meshlet (rebased), for the expand pass
  std::vector<uint32_t> m_tri;   // meshletTriangles of the pool (GeometryPool.ts:430): at firstIndex / 3
  // occluder raster inputs: the pool's vertex positions and meshlet vertex lists (GeometryPool.ts:221-226,
  // 386-429), and per meshlet its list's (offset, vertex count), count 0 for meshes added without positions
  std::vector<float4> m_verts;
  std::vector<uint32_t> m_mverts;
  std::vector<uint2> m_vlist;
  std::vector<uint32_t> m_run;   // LOD runs (see kMaxRun)
  std::vector<uint2> r_span;
  std::vector<float4> r_lod, r_parent;
  std::vector<float2> r_err;
  std::vector<uint32_t> r_block;  // LOD blocks (see kBlockRuns)
  std::vector<uint2> level_blocks;  // per level-table entry pair: the level's first block, last block + 1
  std::vector<uint2> b_span;
  std::vector<float4> b_lod, b_par, b_err;
  uint64_t index_used = 0;  // indices of the meshes added so far
  std::vector<float4> i_col[4];
  std::vector<float> i_slot;  // mesh slot of each instance
  std::vector<float4> c_sphere, c_rec, c_bound;
  bool dirty_table = true, dirty_meshlets = true, dirty_instances = true, dirty_cells = true;
  // Every instance matrix affine (last row 0, 0, 0, 1) with finite entries / every run's LOD spheres finite:
  // together they select the meshlet pass that takes the transformed w as 1.
  bool affine_instances = false, finite_runs = false;
  // device scene
  vgi::DevBuf d_mesh_table, d_cull, d_lod, d_parent, d_err, d_tris, d_col[4], d_csphere, d_crec, d_cbound;
  vgi::DevBuf d_run, d_rspan, d_rlod, d_rparent, d_rerr;
  vgi::DevBuf d_rblock, d_bspan, d_blod, d_bpar, d_berr, d_level_blocks;
  vgi::DevBuf d_first, d_tri, d_islot;  // expand pass inputs
  // Occluder raster inputs: each meshlet's vertex positions in its vertex-list order (x, y, z, 0), so a
  // meshlet's are contiguous; the vertex lists' (offset, count).
  vgi::DevBuf d_vpos, d_vlist;
  vgi::DevBuf d_raster_chunks, d_raster_count;  // the raster's queue of large-triangle rectangles
  vgi::DevBuf d_raster_views;  // the raster records of the last vg_hzb_from_cuts
  // per view slot
  vgi::DevBuf d_counters, d_args, d_metrics, d_draw_ptrs, d_hzb_ptrs, d_index_ptrs;
  std::vector<vgi::DevBuf> d_draws, d_hzbs;
  std::vector<vgi::DevBuf> d_indices;  // 2 per slot (camera, shadow) on contexts with the expand pass
  // two-phase occlusion (allocated by the first view that asks for it): the second phase's counters and
  // args per slot, the recorded rejects (instance bitmap, meshlet records) and the per-slot phase words
  vgi::DevBuf d_counters2, d_args2, d_reject_bits, d_reject_meshlets, d_phase;
  uint32_t reject_words = 0;
  std::vector<uint8_t> phase1_ready;  // per slot: its last cut recorded rejects and no second phase ran since
  std::vector<uint8_t> phase2_ran;    // per slot: a second phase appended to its last cut
  std::vector<uint8_t> hzb_valid;
  std::vector<uint4*> h_draw_ptrs;
  std::vector<float*> h_hzb_ptrs;
  uint32_t counter_stride = 0, args_stride = 0, hzb_floats = 0;
  vgi::DevBuf d_hzb_stage;
  uint64_t layout_version = 1;
  // the scene's own launch lane (the context's stream)
  vgi::Lane lane;
  // profiling: events around each kernel of eager launch sequences on the scene's lane
  bool profile = false;
  std::vector<hipEvent_t> event_pool;
  std::vector<std::vector<hipEvent_t>> pending;  // per sequence: kKernelSlots + 1 events
  double kernel_ms[16] = {0};
  uint64_t profiled_sequences = 0;
  // footprint bitmaps (measurement only)
  vgi::DevBuf d_fp_instances, d_fp_meshlets, d_fp_counts;
  bool footprint = false;
  // launch schedule of the persistent passes
  vg_schedule schedule{6, 10, 5};
  uint64_t device_bytes = 0;
};

namespace vgi {

int fail(vg_ctx* ctx, int code, const std::string& message);
int hip_fail(vg_ctx* ctx, hipError_t e, const char* what);
// Uploads dirty scene buffers and allocates the per-slot outputs of `slot` (context stream, synchronous).
int upload_scene(vg_scene* s);
int ensure_slot(vg_scene* s, uint32_t slot);
// Creates a lane's persistent scratch (view records for max_views, flags, ring); stream: nullptr makes
// one of its own (non-blocking, default priority), else the lane uses `stream`.
int lane_init(vg_ctx* ctx, Lane* lane, hipStream_t stream);
void lane_destroy(Lane* lane);
// Allocates the lane's view records, budget tables and staging ring for max_views (waits for the lane).
int lane_reserve_views(vg_ctx* ctx, Lane* lane);
// Bytes of the transient buffers (queue, work list) a launch of `count` views of `s` needs.
void transient_sizes(const vg_scene* s, uint32_t count, size_t* queue_bytes, size_t* work_items);
// Validates views, uploads the scene, sizes the lane's scratch and stages the view records on the
// lane's stream. `fresh`: every view's controller state starts at rest (scale 1), as on a new slot.
int prepare_cut(vg_scene* s, Lane* lane, const vg_view* views, uint32_t count, bool fresh);
// Enqueues the launch sequence of one cut on the lane's stream.
int enqueue_cut(vg_scene* s, Lane* lane, uint32_t count, bool capturing);
// Replays (capturing first when needed) the launch sequence from the lane's graph.
int launch_graph(vg_scene* s, Lane* lane, uint32_t count);
// Enqueues the second phase of two-phase occlusion for the views staged by prepare_cut.
int enqueue_phase2(vg_scene* s, Lane* lane, uint32_t count);
// Copies a view slot's results out on `stream` and waits for it.
int read_slot(vg_scene* s, hipStream_t stream, uint32_t view_slot, uint32_t* counters, uint32_t* draw_args, uint32_t* draw_list,
              uint32_t max_entries, vg_metrics* metrics);

}  // namespace vgi
