/*
 * vgcut — the per-frame GPU-driven cut of three-virtual-geometry on MI355X (gfx950), as a C ABI.
 *
 * The reference runs this path as TSL compute nodes dispatched by VirtualGeometry.update()
 * (src/core/runtime/VirtualGeometry.ts:392-507) over the storage buffers of a GeometryPool
 * (src/core/runtime/GeometryPool.ts:568-1092). Each entry point below names the reference interface
 * it replaces. Buffers crossing this boundary keep the reference's layouts:
 *   meshlet bounds   16 f32 per meshlet (buildVirtualMesh.ts:85-91)
 *   meshlet info     4 u32 per meshlet  (buildVirtualMesh.ts:92-96)
 *   instance matrix  16 f32 column-major per instance (VirtualMesh.ts:32-37)
 *   counters         GeometryPool.ts:598-640 (12 globals, per (cut, slot, part) count/start, 8 state words)
 *   draw args        5 u32 per (cut, slot, part), drawIndexedIndirect (GeometryPool.ts:605-606, 976-995)
 *   draw list        uvec4 per drawn meshlet (instance, meshlet, firstIndex, triCount | lo << 8 | hi << 16)
 *   HZB              68-float header + mips (OcclusionCulling.ts:30-34, 92-113)
 * Instance ids in the draw list are GPU slots: the library stores each mesh's instances in Morton
 * order (VirtualMesh.ts:171-174); vg_instances_set returns the order.
 *
 * Errors are negative return codes; vg_last_error() describes the last one. Calls on one context are
 * not thread safe and are stream-ordered on the context's stream.
 */
#ifndef VGCUT_H
#define VGCUT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VG_OK 0
#define VG_E_INVALID (-1)
#define VG_E_NOMEM (-2)
#define VG_E_CAPACITY (-3)
#define VG_E_DEVICE (-4)

/* Per-context capacities (VirtualGeometryOptions, VirtualGeometry.ts:62-75, 191-198). */
typedef struct vg_config {
  uint32_t max_triangles;        /* camera triangles per view per frame (default 10M) */
  uint32_t max_shadow_triangles; /* shadow triangles per view per frame (default 4M) */
  uint32_t max_meshlets;         /* drawn meshlets per cut per view (default 1 << 20, at most 1 << 24) */
  uint32_t max_views;            /* view slots: controller state, HZB and outputs are per slot */
  uint32_t mesh_slots;           /* mesh slot capacity of a scene (GeometryPool meshSlotCapacity, default 64) */
  uint32_t hzb_max_size;         /* largest HZB side (OcclusionCulling maxSize, default 1024) */
  uint32_t budget_bins;          /* thresholds of the triangle-budget histogram (default 256) */
  int32_t device;                /* HIP device ordinal; negative: host-only context (pool layout, no cuts) */
  /*
   * 1: every cut ends with the expand pass (GeometryPool.ts:997-1021): each view slot holds an index
   * buffer per cut of max_triangles * 3 (camera) and max_shadow_triangles * 3 (shadow) u32, and every
   * mesh must come with its meshlet triangles. 0: cuts end at the indirect args (no index buffers).
   */
  uint32_t expand;
} vg_config;

/* The VirtualMeshData fields the cut reads (buildVirtualMesh.ts:75-138). */
typedef struct vg_mesh_desc {
  uint32_t meshlet_count;
  const float* meshlet_bounds;      /* 16 per meshlet */
  const uint32_t* meshlet_info;     /* 4 per meshlet */
  uint32_t level_count;
  const uint32_t* level_ranges;     /* 2 per level */
  const float* level_errors;        /* 2 per level */
  const float* level_guard_errors;  /* 2 per level */
  float bounding_sphere[4];
  float lod_bounds_sphere[4];
  float cull_bounds_sphere[4];
  /*
   * VirtualMeshData.meshletTriangles (buildVirtualMesh.ts:75-138, GeometryPool.ts:430): 3 local vertex
   * ids per triangle packed a | b << 8 | c << 16; meshlet i's triangles start at meshlet_info[4i + 1] / 3.
   * Needed by the expand pass; may be NULL (triangle_count 0) on a context without it.
   */
  const uint32_t* meshlet_triangles;
  uint64_t triangle_count;
  /*
   * VirtualMeshData.positions (xyz per vertex) and meshletVertices (the mesh's vertex ids per meshlet, at
   * meshlet_info[4i + 3]; buildVirtualMesh.ts:75-138), stored as the pool's vertexData / meshletVertices
   * (GeometryPool.ts:221-226, 386-429). Needed by the occluder raster (vg_hzb_from_cut) only, together with
   * meshlet_triangles; may be NULL (counts 0): the mesh then draws nothing into occluder depth.
   */
  const float* positions;
  uint32_t vertex_count;
  const uint32_t* meshlet_vertices;
  uint64_t meshlet_vertex_count;
} vg_mesh_desc;

/*
 * One view of a cut: the uniforms VirtualGeometry.update() sets (VirtualGeometry.ts:78-183, 400-437).
 * `slot` picks the view slot whose controller state, HZB and outputs this view uses.
 */
typedef struct vg_view {
  float view_matrix[16];        /* camera.matrixWorldInverse, column-major */
  float projection_matrix[16];  /* unjittered projection, column-major (occlusion test) */
  float frustum_planes[24];     /* 6 x (normal xyz, constant), Frustum.setFromProjectionMatrix(P * V) */
  float shadow_sweep[3];
  float proj_scale;             /* projection[5] * viewport height / 2 */
  float near_plane;
  float error_threshold;        /* pixels */
  float shadow_error_scale;     /* default 3 */
  float lod_blend_max;          /* max(1, lodBlend), default 2 */
  float lod_scale_limit;        /* max(1, maxErrorThreshold / errorThreshold) */
  float min_pixel_radius;       /* default 0.7 */
  uint32_t frustum_culling;     /* 1: on */
  uint32_t occlusion;           /* 1: occlusion culling wanted this frame (needs a valid HZB in the slot);
                                   2: the same, in two phases (see vg_cut_phase2) */
  uint32_t slot;                /* view slot, < max_views */
  uint32_t triangle_budget;     /* 0: off; else the error threshold is chosen on the GPU (see vg_cut) */
  float threshold_range[2];     /* budget: smallest and largest threshold (VirtualGeometry.thresholdRange) */
  float max_error_threshold;    /* budget: maxErrorThreshold (VirtualGeometry.ts:124) */
} vg_view;

/* Work metrics of the last cut of a view slot (each per cut: camera, shadow). */
typedef struct vg_metrics {
  uint64_t cells_tested;
  uint64_t instances_tested;
  uint64_t work_items[2];
  uint64_t clusters_tested[2];  /* meshlet-pass range lengths + fast-path meshlets, retries included */
  uint64_t occlusion_tests;
  float error_threshold;        /* threshold the cut used (the budget's pick in budget mode) */
  uint32_t budget_bin;          /* budget mode: chosen bin */
  uint64_t budget_cluster_tests; /* budget mode: clusters binned by the histogram pass */
} vg_metrics;

typedef struct vg_ctx vg_ctx;
typedef struct vg_scene vg_scene;

/* VirtualGeometry constructor (VirtualGeometry.ts:191-198). NULL config: defaults. */
int vg_create(const vg_config* config, vg_ctx** out);
void vg_destroy(vg_ctx* ctx);
const char* vg_last_error(const vg_ctx* ctx);

/* A resident scene = one GeometryPool (GeometryPool.ts:260-280). */
int vg_scene_create(vg_ctx* ctx, vg_scene** out);
void vg_scene_destroy(vg_scene* scene);
/* Device bytes a scene holds (admission by HBM footprint). */
uint64_t vg_scene_bytes(const vg_scene* scene);

/* GeometryPool.add geometry part (GeometryPool.ts:302-355, 386-460): returns the mesh slot. */
int vg_mesh_add(vg_scene* scene, const vg_mesh_desc* mesh, uint32_t* slot);
/*
 * Instances of a mesh (GeometryPool.writeInstances + computeCell, GeometryPool.ts:484-545; Morton order
 * VirtualMesh.ts:171-174, GeometryPool.ts:1263-1311). `matrices`: 16 f32 per instance; `colors`: 4 f32
 * per instance or NULL. `gpu_to_input` (n entries, may be NULL) receives the caller instance of each slot.
 * A new count moves the mesh to a fresh instance range; the columns are repacked (every mesh's range moves,
 * in slot order) once more than half of them would be dead.
 */
int vg_instances_set(vg_scene* scene, uint32_t slot, const float* matrices, const float* colors, uint32_t count,
                     uint32_t* gpu_to_input);
/*
 * GeometryPool.remove (GeometryPool.ts:357-363), as VirtualGeometry.unregister calls it (VirtualGeometry.ts:370-376):
 * the slot is disabled and freed for the next vg_mesh_add. Its instances and cells are dropped (the reference
 * leaves its cells pointing at the disabled slot, which a later add into the slot would enable again); the
 * instance columns are repacked once more than half of them are dead. Its geometry stays in the pool until
 * the scene is destroyed, as the reference keeps it until a repack it does not implement.
 */
int vg_mesh_remove(vg_scene* scene, uint32_t slot);
/* GeometryPool.writeFlags (GeometryPool.ts:463-477). max_draw_distance >= 1e30: none. */
int vg_mesh_flags(vg_scene* scene, uint32_t slot, float max_draw_distance, float min_pixel_radius, uint32_t enabled,
                  uint32_t shadow, uint32_t occlusion);

/* OcclusionCulling HZB copy + reduce (OcclusionCulling.ts:275-352) from a host depth image (row 0 at top). */
int vg_hzb_set(vg_scene* scene, uint32_t view_slot, const float* depth, uint32_t width, uint32_t height);
int vg_hzb_invalidate(vg_scene* scene, uint32_t view_slot);
/*
 * The occluder pass (OcclusionCulling.ts:4-21, 244-273) as a software raster on the device: the camera
 * entries of the view slot's last cut (draw-list positions [0, drawn); with `with_phase2`, also what the
 * slot's second phase appended, when it ran after that cut) are drawn depth-only with `view_proj`
 * (column-major clip = view_proj * world, WebGPU depth in [0, 1]) into a width x height depth image
 * (cleared to 1, row 0 at the top), and the slot's HZB is built from it, in stream order with the cut
 * (no host round trip; vg_hzb_set builds it from a host image instead). Meshes added without positions
 * draw nothing. Triangles are skipped when back-facing, degenerate, with a vertex in front of the near
 * plane (clip z < 0 or w <= 0) or more than 2^20 pixels off screen: fewer occluders only make the depth
 * farther, never hide what is visible. Coverage: pixel centres on or inside all three edges of the
 * triangle snapped to 1/256 pixel; depth: the barycentric blend of the vertices' clip z / w, minimum kept.
 */
int vg_hzb_from_cut(vg_scene* scene, uint32_t view_slot, const float* view_proj, uint32_t width, uint32_t height,
                    uint32_t with_phase2);
/* One raster of vg_hzb_from_cuts: the arguments of a vg_hzb_from_cut. */
typedef struct vg_raster {
  uint32_t view_slot;
  uint32_t width, height;
  uint32_t with_phase2;
  float view_proj[16];
} vg_raster;
/* vg_hzb_from_cut for `count` distinct view slots (at most 256) in one launch sequence: the same depth
 * images and HZBs as `count` vg_hzb_from_cut calls. The reference's occluder pass draws one camera at a
 * time (OcclusionCulling.ts:244-273); this batched form has no counterpart there. */
int vg_hzb_from_cuts(vg_scene* scene, const vg_raster* rasters, uint32_t count);
/* The slot's HZB buffer (OcclusionCulling hzbAttribute): `size_floats` receives its length; dst may be NULL.
 * Each level has a region sized for hzb_max_size; its texels outside the level's width x height are 0. */
int vg_hzb_read(vg_scene* scene, uint32_t view_slot, float* dst, uint64_t capacity_floats, uint64_t* size_floats);

/*
 * The compute part of VirtualGeometry.update() (VirtualGeometry.ts:462-491): reset, cells, instances,
 * meshlets, prefix, capacity retry rounds and steer for `count` views in one launch sequence, enqueued
 * on the context's stream. Views must use distinct slots. With `triangle_budget` > 0 a view's error
 * threshold is the smallest of `budget_bins` log-spaced thresholds over threshold_range whose plain cut
 * has at most that many triangles (projected-error histogram), else the largest.
 */
int vg_cut(vg_scene* scene, const vg_view* views, uint32_t count);
/* Same as vg_cut, replayed from a captured HIP graph when the view count matches the last capture. */
int vg_cut_graph(vg_scene* scene, const vg_view* views, uint32_t count);
/*
 * vg_cut without allocating: no counterpart in the reference. For callers that cannot allocate device
 * memory while they run (code being captured into a HIP graph, tight serving loops, hosts that forbid
 * hipMalloc in their frame loop). The cut's launch buffers (view records, the (view, cell) queue, the
 * work list, its counters and flags, budget tables) live in a static array of 64 MiB that the library
 * holds on each device it runs on until the process ends, instead of buffers the scene's lane grows.
 * Limits: at most 64 views; the (view, cell) queue (8 bytes per view per cell) must fit, and the work list
 * gets what is left (28 bytes per item): a launch whose work list overflows returns VG_E_CAPACITY;
 * contexts with the expand pass are refused. The scene must already be on the device and each view slot
 * must have outputs, which a vg_cut of those slots after the last scene change arranges (otherwise
 * VG_E_INVALID, nothing launched). Calls no hipMalloc/hipFree; waits for the cut before returning.
 * Every call on a device shares that device's array: calls that overlap on one device (from two threads,
 * two contexts) race and must not be made.
 */
int vg_cut_static(vg_scene* scene, const vg_view* views, uint32_t count);

/*
 * Two-phase occlusion (the reference tests occluders in one phase, OcclusionCulling.ts:4-21; the second
 * phase is this library's): a vg_cut of views with occlusion = 2 records its camera cut's occlusion
 * rejects (instances and meshlets of the final round). After the caller has drawn that cut and set each
 * slot's new HZB (vg_hzb_set, from the depth it drew), vg_cut_phase2 with the same views re-tests the
 * rejects against it: rejected instances go through the instance and meshlet passes again, rejected
 * meshlets through the occlusion test, at the cut's final band. What is visible now is appended to the
 * camera draw list after the cut's entries, with its own regions laid out after the cut's requested
 * indices and its own draw args (vg_read_phase2); with the expand pass, its indices are written there.
 * It must directly follow the views' vg_cut on the scene (no instance changes in between).
 */
int vg_cut_phase2(vg_scene* scene, const vg_view* views, uint32_t count);
/* The second phase's counters (reference layout; C_SELECTED, C_DRAWN and C_REQUESTED count both phases,
   the per-mesh (count, start) pairs and args are the second phase's) and draw args, after vg_sync. */
int vg_read_phase2(vg_scene* scene, uint32_t view_slot, uint32_t* counters, uint32_t* draw_args);
/* Meshlet rejects the slot's last cut recorded, and whether more were dropped (more than max_meshlets). */
int vg_phase_rejects(vg_scene* scene, uint32_t view_slot, uint32_t* meshlets, uint32_t* dropped);
/* Waits for the context's stream. */
int vg_sync(vg_scene* scene);

/*
 * Reads a view slot's results after vg_sync (GeometryPool.readCounters, GeometryPool.ts:1135-1151):
 * counters: 12 + 8 * mesh_slots + 8 words; draw_args: 20 * mesh_slots words; draw_list: up to
 * `max_entries` uvec4 per cut (camera entries, then shadow entries at offset max_entries). Any pointer
 * may be NULL.
 */
int vg_read(vg_scene* scene, uint32_t view_slot, uint32_t* counters, uint32_t* draw_args, uint32_t* draw_list,
            uint32_t max_entries, vg_metrics* metrics);
/*
 * The index buffer of a view slot's cut (0 camera, 1 shadow) that the expand pass wrote
 * (GeometryPool.ts:997-1021; cameraIndex / shadowIndex, read by the indirect draws, VirtualMesh.ts:266-284):
 * `count` u32 from index `first`, after vg_sync. An index is (draw-list position << 7) | local vertex,
 * where the position counts shadow entries from max_meshlets. Contexts created with expand only.
 */
int vg_index_read(vg_scene* scene, uint32_t view_slot, uint32_t cut, uint64_t first, uint64_t count, uint32_t* dst);
/* Device address and capacity (u32 indices) of that index buffer, for consumers on the same device. */
int vg_index_buffer(vg_scene* scene, uint32_t view_slot, uint32_t cut, void** device_ptr, uint64_t* capacity);
/* Capacity-controller state words of a view slot (8 u32, x10000). */
int vg_state_get(vg_scene* scene, uint32_t view_slot, uint32_t* state);
int vg_state_set(vg_scene* scene, uint32_t view_slot, const uint32_t* state);

/*
 * The pool's host-side buffers in the reference layout (GeometryPool.ts:221-231, the typed arrays behind
 * its GrowBuffers): what a debugger or CPU consumer of the pool reads. `size_bytes` receives the size;
 * `dst` may be NULL to query it. Instances are in GPU (Morton) order: 4 matrix columns, then
 * (tint rgb = 1, mesh slot).
 */
#define VG_POOL_MESH_TABLE 0     /* f32: 24 per mesh slot, then the level tables (8 per level) */
#define VG_POOL_INSTANCES 1      /* f32: 20 per instance */
#define VG_POOL_CELLS 2          /* f32: 8 per cell */
#define VG_POOL_MESHLET_BOUNDS 3 /* f32: 16 per meshlet (pool-global ids) */
#define VG_POOL_MESHLET_INFO 4   /* u32: 4 per meshlet, firstIndex rebased to the pool */
#define VG_POOL_MESHLET_TRIANGLES 5 /* u32: 1 per triangle (packed local ids), at firstIndex / 3 */
#define VG_POOL_VERTICES 6       /* f32: 4 per vertex (x, y, z, 0), pool-global ids */
#define VG_POOL_MESHLET_VERTICES 7 /* u32: pool-global vertex ids; meshletInfo[3] is each meshlet's offset */
int vg_scene_pool_read(const vg_scene* scene, uint32_t buffer, void* dst, uint64_t capacity_bytes, uint64_t* size_bytes);
/* Highest used mesh slot + 1 (GeometryPool.meshSlots). */
uint32_t vg_scene_mesh_slots(const vg_scene* scene);
/* First GPU instance slot and instance count of a mesh (PoolEntry.instanceBase/instanceCount, GeometryPool.ts:161-162). */
int vg_mesh_instance_base(const vg_scene* scene, uint32_t slot, uint32_t* base, uint32_t* count);

/*
 * Measurement. vg_profile(scene, 1) times every kernel of eager launch sequences with HIP events on the
 * context's stream, and the raster kernels of vg_hzb_from_cut; vg_profile_read returns the summed
 * milliseconds per kernel position (13 positions, `names` receives their names), the number of cut
 * launch sequences, and resets them. vg_footprint runs one cut
 * (like vg_cut, then waits) and counts the distinct instances tested and distinct meshlet records in
 * candidate ranges over all views (the unique inputs one launch must read).
 */
int vg_profile(vg_scene* scene, uint32_t enable);
int vg_profile_read(vg_scene* scene, double* kernel_ms, uint32_t max_kernels, uint64_t* sequences, const char** names);
int vg_footprint(vg_scene* scene, const vg_view* views, uint32_t count, uint64_t* instances, uint64_t* meshlets);

/*
 * Launch schedule of the persistent passes. The batched form has no counterpart in the reference
 * (which dispatches one workgroup per work item, GeometryPool.ts:926-968); results do not depend on the
 * schedule, only the time does. instance_blocks_per_cu: instance-pass workgroups (256 threads) per CU,
 * 1-16. meshlet_blocks_per_cu: meshlet-pass workgroups per CU, 1-32. meshlet_waves_per_simd: the
 * occupancy the meshlet pass is compiled for (VG_MESHLET_VARIANTS: 4, 5, 6 or 7; higher = fewer registers).
 * vg_tune times candidate schedules on `count` views (`repeats` eager cuts each, the median kept, the
 * controller state restored after each), sets the fastest on the scene and returns it with its time.
 */
typedef struct vg_schedule {
  uint32_t instance_blocks_per_cu;
  uint32_t meshlet_blocks_per_cu;
  uint32_t meshlet_waves_per_simd;
} vg_schedule;
int vg_schedule_get(const vg_scene* scene, vg_schedule* schedule);
int vg_schedule_set(vg_scene* scene, const vg_schedule* schedule);
int vg_tune(vg_scene* scene, const vg_view* views, uint32_t count, uint32_t repeats, vg_schedule* best, float* best_ms);

/*
 * Request server: cut requests against resident scenes, batched continuously (no counterpart in the
 * reference, which cuts its views inside VirtualGeometry.update, VirtualGeometry.ts:392-507; this is
 * the serving form of that path). A worker thread keeps a queue of requests; whenever a lane (a HIP
 * stream) is free it launches one batch: queued requests of one scene that fit the batch's view limit,
 * the scene's free view slots and the HBM budget. Under VG_SERVE_FIFO the batch starts at the oldest
 * request that fits and takes the scene's others oldest first; under VG_SERVE_SHORTEST (the default)
 * it starts at the request of least estimated run time (views x the scene's measured time per view)
 * and takes the scene's others cheapest first, except that a request queued longer than max_bypass_us
 * starts the next batch, oldest first. A batch's transient buffers
 * (cell queue, work list) come from the device's stream-ordered memory pool and are returned when it
 * ends; the work list starts at first_work_items and a batch that overflows it runs again at its worst-case
 * size (every (view, instance, cut)), outside the budget check. Requests that do not fit wait while
 * others overtake them, up to max_bypass_us (default 2000 under FIFO, 20000 under SHORTEST). Batches on
 * different lanes overlap on the device. Each request's views start from a controller at rest (scale 1),
 * so a result does not depend on the batch it ran in. Results stay in the request's view slots until
 * vg_server_release. Scenes given to a server must not be changed or cut directly while it runs, and
 * only the vg_server_* calls may be made on its context meanwhile.
 */
typedef struct vg_server_config {
  uint64_t hbm_budget;          /* bytes for resident scenes + in-flight batches; 0: 90% of device memory */
  uint32_t lanes;               /* batches in flight, 1-16 (default 4) */
  uint32_t max_batch_views;     /* views per batch, 1 - max_views (default max_views) */
  uint32_t max_batch_requests;  /* requests per batch; 0: no limit, 1: one request per launch */
  uint32_t max_bypass_us;       /* how long later requests may overtake one that does not fit (default 2000) */
  uint32_t first_work_items;    /* work-list items of a batch's first launch (default 1 << 22) */
  uint32_t policy;              /* VG_SERVE_SHORTEST (default) or VG_SERVE_FIFO */
  uint32_t max_blocks_per_cu;   /* cap on the persistent blocks per CU of a batch's launches, so batches on
                                   other lanes find room beside it (0: the scene's tuned schedule) */
} vg_server_config;

enum { VG_SERVE_SHORTEST = 0, VG_SERVE_FIFO = 1 };

typedef struct vg_server vg_server;

typedef struct vg_request_done {
  uint64_t ticket;
  int32_t status;           /* VG_OK or a negative code */
  uint32_t views;
  uint32_t batch_requests;  /* requests in the batch it ran in */
  uint32_t lane;
  double queued_us;         /* submit -> launch */
  double run_us;            /* launch -> completion seen by the worker */
  double total_us;          /* submit -> completion */
} vg_request_done;

typedef struct vg_server_stats {
  /* inflight_bytes: the lanes' transient buffers (cell queues, work lists), kept between batches;
     peak_inflight_bytes: the most they held. resident_bytes + inflight_bytes <= hbm_budget. */
  uint64_t hbm_budget, resident_bytes, inflight_bytes, peak_inflight_bytes;
  uint64_t batches, requests_done, max_batch_requests_seen, reruns;
  uint32_t queued, running;
} vg_server_stats;

int vg_server_create(vg_ctx* ctx, const vg_server_config* config, vg_server** out);
/* Stops the worker after the batches in flight end; queued requests are dropped. */
void vg_server_destroy(vg_server* server);
const char* vg_server_last_error(const vg_server* server);
/* Makes a scene resident for serving: uploads it and allocates all its view slots; refused
   (VG_E_NOMEM) when its bytes do not fit the budget next to the scenes already resident. */
int vg_server_add_scene(vg_server* server, vg_scene* scene, uint32_t* scene_id);
/* Queues a request of `count` views (their `slot` fields are ignored: the server assigns slots). */
int vg_server_submit(vg_server* server, uint32_t scene_id, const vg_view* views, uint32_t count, uint64_t* ticket);
/* Completed requests, oldest first; waits up to wait_us for one when none is ready. */
int vg_server_poll(vg_server* server, vg_request_done* done, uint32_t max_done, uint32_t* count, uint32_t wait_us);
/* Results of view `view_index` of a completed request (as vg_read). */
int vg_server_read(vg_server* server, uint64_t ticket, uint32_t view_index, uint32_t* counters, uint32_t* draw_args,
                   uint32_t* draw_list, uint32_t max_entries, vg_metrics* metrics);
/* Frees a completed request's view slots. */
int vg_server_release(vg_server* server, uint64_t ticket);
int vg_server_stats_get(const vg_server* server, vg_server_stats* stats);

/*
 * The triangle budget's threshold table: bins values lo * (hi / lo)^(b / (bins - 1)) rounded to f32 (the
 * last one exactly hi). A view's histogram bins clusters by these values (times the controller scale).
 */
int vg_budget_thresholds(float lo, float hi, uint32_t bins, float* out);

/* Layout helpers: counter words and draw-args words for the context's mesh slot capacity. */
uint32_t vg_counter_words(const vg_ctx* ctx);
uint32_t vg_draw_args_words(const vg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
