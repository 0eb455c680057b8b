/*
 * ORACLE — test infrastructure only. A sequential CPU restatement of the reference's per-frame GPU
 * cut (src/core/runtime/GeometryPool.ts:568-1092 and OcclusionCulling.ts:119-170, 275-352) in
 * binary32 arithmetic, in the reference kernels' expression order, with no FMA contraction and
 * correctly rounded division and square root. Only tests, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it; nothing in the product links it.
 *
 * Buffers use the reference's layouts (GeometryPool.ts:598-640): counters = [0] visible cells,
 * [1] occluded instances, [2 + cut*5 + {SELECTED, DRAWN, SURVIVED, WORK, REQUESTED}], then per
 * (cut, slot, part) the (count, start) pair at 12 + cut*slots*4 + slot*4 + part*2, then the 8
 * fixed-point controller words (x10000).
 */
#ifndef VGORACLE_H
#define VGORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene {
  uint32_t slots;            /* mesh slot capacity: stride of counters and draw args */
  uint32_t mesh_slots;       /* highest used slot + 1 */
  const float* mesh_table;   /* 24 floats per slot (GeometryPool.ts:449-460), level tables after */
  const float* instances;    /* 20 floats per instance (GeometryPool.ts:484-497) */
  uint32_t instance_count;
  const float* cells;        /* 8 floats per cell (GeometryPool.ts:519-545) */
  uint32_t cell_count;
  const float* bounds;       /* 16 floats per meshlet (pool-global ids) */
  const uint32_t* info;      /* 4 u32 per meshlet */
  uint32_t meshlet_count;
  uint32_t max_meshlets;     /* draw capacity per cut (maxDrawn) */
  uint32_t cut_capacity[2];  /* index capacity per cut: camera, shadow (triangles * 3) */
} or_scene;

typedef struct or_view {
  float view[16];            /* column-major world -> view */
  float proj[16];            /* column-major projection (occlusion test) */
  float planes[24];          /* 6 x (nx, ny, nz, w) */
  float sweep[3];            /* shadowSweep */
  float proj_scale;
  float near_plane;
  float error_threshold;
  float shadow_error_scale;
  float lod_blend_max;
  float lod_scale_limit;
  float min_pixel_radius;
  uint32_t frustum_culling;
  uint32_t occlusion_active; /* the HZB holds occluders this frame (OcclusionCulling.activeNode) */
  const float* hzb;          /* HZB buffer: 68-float header + mips (OcclusionCulling.ts:30-34, 284-300) */
} or_view;

typedef struct or_stats {
  uint64_t cells_tested;
  uint64_t instances_tested;     /* instances of visible cells, per pass (retries included) */
  uint64_t work_items[2];
  uint64_t clusters_tested[2];   /* meshlet-pass range lengths + fast-path meshlets */
  uint64_t occlusion_tests;
  uint64_t retries[2];
  /* rounds (checks and the steer) whose cut overflowed the draw list: the index counts then depend on the
     order of the appends (GeometryPool.ts:957-962 counts only appends that got a slot), and with them the
     fill and the controller state */
  uint64_t draw_overflow[2];
  /* fp64 check of the LOD decision (cut.ts:26-48 form) on every tested cluster */
  uint64_t lod_checked;
  uint64_t near_threshold;       /* within 1e-6 relative of a threshold in fp64 */
  uint64_t mismatch_near;        /* fp32 and fp64 decisions differ, flagged near-threshold */
  uint64_t mismatch_hard;        /* fp32 and fp64 decisions differ, not near a threshold */
} or_stats;

/* Counter words for `slots` mesh slots (without the 8 controller words). */
uint32_t or_counter_words(uint32_t slots);

/*
 * One frame of the cut (reset, cells, instances, meshlets, prefix, 2 retry rounds, steer).
 * `state` holds the 8 controller words and is updated in place; `counters` receives
 * or_counter_words(slots) words, `draw_args` 2*slots*2*5 words, `draw_list` 2*max_meshlets*4 words
 * (entries at or beyond the selected count of a cut are left as they were).
 * `fp64_check` != 0 compares every tested cluster's LOD decision with the fp64 form.
 */
int or_cut(const or_scene* scene, const or_view* view, uint32_t* state, uint32_t* counters, uint32_t* draw_args,
           uint32_t* draw_list, int fp64_check, or_stats* stats);

/*
 * Two-phase occlusion (the library's design; the reference has one phase): or_cut of view1 recording the
 * camera cut's occlusion rejects of its final round (instances; meshlets up to reject_capacity), then the
 * second phase against view2's HZB, appended to draw_list after the cut's entries, into counters2/args2
 * (C_SELECTED/C_DRAWN/C_REQUESTED over both phases, per-mesh pairs and args of the second phase).
 * rejects (may be NULL): [0] instances rejected, [1] meshlets rejected, [2] whether meshlet records were
 * dropped.
 */
int or_cut_two_phase(const or_scene* scene, const or_view* view1, const or_view* view2, uint32_t* state, uint32_t* counters1,
                     uint32_t* args1, uint32_t* counters2, uint32_t* args2, uint32_t* draw_list, uint32_t reject_capacity,
                     or_stats* stats1, or_stats* stats2, uint32_t* rejects);

/* HZB build (OcclusionCulling.ts:275-352): header + mip 0 = depth + max pyramid. Returns floats written. */
uint32_t or_hzb_size(uint32_t max_size);
int or_hzb_build(const float* depth, uint32_t width, uint32_t height, uint32_t max_size, float* hzb);

/* occluded() of OcclusionCulling.ts:119-170 on a view-space sphere. */
int or_occluded(const float* proj, float near_plane, const float* hzb, const float* view_center, float radius);

/* The fp32 natural logarithm both sides use for blend steps. */
float or_logf(float x);

/*
 * Projected error histogram for the triangle budget (our design, not in the reference): over the
 * clusters an instance pass at band [lo, hi] would test, bin b of `thresholds` (ascending) gets the
 * triangles of clusters in the plain cut at thresholds[b]. Returns 0 on success.
 */
int or_budget_histogram(const or_scene* scene, const or_view* view, const uint32_t* state, const float* thresholds,
                        uint32_t bins, uint64_t* triangles);

/*
 * The expand pass (GeometryPool.ts:997-1021) over a given cut result: `counters` and `draw_list` as
 * or_cut or the GPU leave them (2 * max_meshlets entries), `triangles` the pool's meshletTriangles.
 * Expands the entries [first, C_DRAWN) (first > 0: the second phase of two-phase occlusion, with its
 * counters). Writes into `indices` (capacity cut_capacity[cut]) and leaves other positions as they are;
 * `written` receives the number of indices written. Returns 0, or < 0 on an entry outside the pool.
 */
int or_expand(const or_scene* scene, const uint32_t* triangles, uint64_t triangle_words, int cut, const uint32_t* counters,
              const uint32_t* draw_list, uint32_t first, uint32_t* indices, uint64_t* written);

/*
 * The occluder pass (OcclusionCulling.ts:4-21, 244-273) as vg_hzb_from_cut rasterizes it: the camera entries
 * draw_list[0 .. count) (uvec4) of a cut, over the pool's instances (20 f32 each), meshlet info (4 u32 each,
 * [3] = offset into meshlet_vertices or ~0 for none), meshletTriangles, meshlet_vertices (pool-global ids)
 * and vertices (4 f32 each), are drawn depth-only with `view_proj` into `depth` (width x height, row 0 at
 * the top), which must hold 1.0 on entry. Returns 0, or -1 on an entry outside the pool.
 */
int or_raster(const float* instances, uint32_t instance_count, const uint32_t* info, uint32_t meshlet_count,
              const uint32_t* triangles, uint64_t triangle_words, const uint32_t* meshlet_vertices,
              uint64_t meshlet_vertex_words, const float* vertices, uint64_t vertex_count, const uint32_t* draw_list,
              uint32_t count, const float* view_proj, uint32_t width, uint32_t height, float* depth);

/*
 * Test hooks. or_test_controller: checkNode (steer = 0) or steerNode (steer = 1) of both cuts on given
 * requested indices, selected meshlets and capacities (GeometryPool.ts:1027-1069), updating `state` (8
 * words); redo[cut] receives whether checkNode redoes that cut. or_test_prefix: prefixPass of one cut
 * (GeometryPool.ts:973-995) on given index counts (2 per slot, slot-major) from region `offset`, into
 * `counters` (or_counter_words(slots) words) and `args` (2 * slots * 2 * 5 words).
 */
int or_test_controller(int steer, const uint32_t* requested, const uint32_t* selected, const uint32_t* capacity,
                       uint32_t max_meshlets, float lod_scale_limit, uint32_t* state, int* redo);
int or_test_prefix(uint32_t slots, uint32_t mesh_slots, int cut, const uint32_t* counts, uint32_t capacity, uint32_t selected,
                   uint32_t max_meshlets, uint32_t offset, uint32_t* counters, uint32_t* args);

#ifdef __cplusplus
}
#endif
#endif
