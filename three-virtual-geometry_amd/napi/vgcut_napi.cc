// vgcut_napi.cc: the N-API addon the TypeScript host loads in place of the TSL compute nodes (build.sh or
// node-gyp with binding.gyp; links lib/libvgcut.so and lib/libvgdag.so). JS values in, C ABI calls out.
#include <node_api.h>
#include <cstdio>
#include <cstring>
#include <vector>
#include "vgcut.h"
#include "vgdag.h"

#define CHECK_NAPI(call)                                              \
  do {                                                                \
    if ((call) != napi_ok) {                                          \
      napi_throw_error(env, nullptr, "vgcut: N-API call failed");     \
      return nullptr;                                                 \
    }                                                                 \
  } while (0)

static napi_value throw_vg(napi_env env, int rc, vg_ctx* ctx) {
  char code[16];
  snprintf(code, sizeof(code), "%d", rc);
  napi_throw_error(env, code, ctx ? vg_last_error(ctx) : "vgcut error");
  return nullptr;
}

template <typename T>
struct ArrayType;
template <>
struct ArrayType<float> {
  static constexpr napi_typedarray_type value = napi_float32_array;
  static constexpr const char* name = "Float32Array";
};
template <>
struct ArrayType<uint32_t> {
  static constexpr napi_typedarray_type value = napi_uint32_array;
  static constexpr const char* name = "Uint32Array";
};

// The data and element count of a typed array of T's own type (Float32Array for float, Uint32Array for
// uint32_t). Absent (undefined or null): nullptr, *n = 0, nothing thrown. Any other value, a typed array
// of another element type included: a TypeError is thrown and nullptr returned with *bad set, so a short
// Uint8Array can never pass a check made in elements of T.
template <typename T>
static T* typed(napi_env env, napi_value v, size_t* n, bool* bad, const char* what) {
  *n = 0;
  *bad = false;
  napi_valuetype vt;
  if (!v || napi_typeof(env, v, &vt) != napi_ok || vt == napi_undefined || vt == napi_null) return nullptr;
  napi_typedarray_type type;
  void* data = nullptr;
  napi_value buffer;
  size_t offset = 0, count = 0;
  bool is_typed = false;
  if (napi_is_typedarray(env, v, &is_typed) != napi_ok || !is_typed ||
      napi_get_typedarray_info(env, v, &type, &count, &data, &buffer, &offset) != napi_ok || type != ArrayType<T>::value) {
    char msg[160];
    snprintf(msg, sizeof(msg), "vgcut: %s must be a %s", what, ArrayType<T>::name);
    napi_throw_type_error(env, nullptr, msg);
    *bad = true;
    return nullptr;
  }
  *n = count;
  return static_cast<T*>(data);
}

static napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v = nullptr;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

// A context and whether a cut on it is running on a worker thread (set and cleared on the JS thread).
struct Ctx {
  vg_ctx* ctx;
  bool in_flight;
};

// createContext({maxTriangles, maxShadowTriangles, maxMeshlets, maxViews, meshSlots, device, expand}) -> external
static napi_value CreateContext(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vg_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  auto u32 = [&](const char* name, uint32_t* out) {
    uint32_t x;
    if (napi_get_value_uint32(env, prop(env, argv[0], name), &x) == napi_ok) *out = x;
  };
  u32("maxTriangles", &cfg.max_triangles);
  u32("maxShadowTriangles", &cfg.max_shadow_triangles);
  u32("maxMeshlets", &cfg.max_meshlets);
  u32("maxViews", &cfg.max_views);
  u32("meshSlots", &cfg.mesh_slots);
  bool expand = false;
  if (napi_get_value_bool(env, prop(env, argv[0], "expand"), &expand) == napi_ok) cfg.expand = expand ? 1u : 0u;
  int32_t device = 0;
  napi_get_value_int32(env, prop(env, argv[0], "device"), &device);
  cfg.device = device;
  vg_ctx* ctx = nullptr;
  if (int rc = vg_create(&cfg, &ctx)) return throw_vg(env, rc, nullptr);
  auto* c = new Ctx{ctx, false};
  napi_value out;
  CHECK_NAPI(napi_create_external(env, c, [](napi_env, void* p, void*) {
    auto* c = static_cast<Ctx*>(p);
    vg_destroy(c->ctx);
    delete c;
  }, nullptr, &out));
  return out;
}

// A scene keeps its context's external alive (ctx_ref) while it exists.
struct Scene {
  Ctx* c;
  vg_ctx* ctx;
  vg_scene* scene;
  napi_ref ctx_ref;
};

// Calls on a context are made from the JS thread only and never while a cut of that context runs on a
// worker thread (the library's calls on one context are not thread safe, and the cut may be capturing a
// graph on the context's stream): every entry point checks the flag and throws instead.
static bool busy(napi_env env, const Ctx* c) {
  if (!c->in_flight) return false;
  napi_throw_error(env, "VG_BUSY", "vgcut: a cut on this context is still running; await its promise first");
  return true;
}

static Scene* scene_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "vgcut: expected a scene");
    return nullptr;
  }
  auto* s = static_cast<Scene*>(p);
  return busy(env, s->c) ? nullptr : s;
}

static Ctx* ctx_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "vgcut: expected a context");
    return nullptr;
  }
  return static_cast<Ctx*>(p);
}

// createScene(ctx) -> external (one GeometryPool)
static napi_value CreateScene(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Ctx* c = ctx_of(env, argv[0]);
  if (!c || busy(env, c)) return nullptr;
  auto* s = new Scene{c, c->ctx, nullptr, nullptr};
  if (int rc = vg_scene_create(s->ctx, &s->scene)) {
    delete s;
    return throw_vg(env, rc, c->ctx);
  }
  if (napi_create_reference(env, argv[0], 1, &s->ctx_ref) != napi_ok) {
    vg_scene_destroy(s->scene);
    delete s;
    napi_throw_error(env, nullptr, "vgcut: N-API call failed");
    return nullptr;
  }
  napi_value out;
  CHECK_NAPI(napi_create_external(env, s, [](napi_env env, void* p, void*) {
    auto* sc = static_cast<Scene*>(p);
    vg_scene_destroy(sc->scene);
    napi_delete_reference(env, sc->ctx_ref);
    delete sc;
  }, nullptr, &out));
  return out;
}

// addMesh(scene, data: VirtualMeshData) -> slot (GeometryPool.add, geometry part; positions and
// meshletVertices, when present, feed the occluder raster)
static napi_value AddMesh(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  size_t nb = 0, ni = 0, nr = 0, ne = 0, ng = 0, n4 = 0;
  bool bad = false;
  vg_mesh_desc d;
  memset(&d, 0, sizeof(d));
#define VG_FIELD(T, dst, count, name)                                      \
  do {                                                                     \
    dst = typed<T>(env, prop(env, argv[1], name), &count, &bad, name);     \
    if (bad) return nullptr;                                               \
  } while (0)
  VG_FIELD(float, d.meshlet_bounds, nb, "meshletBounds");
  VG_FIELD(uint32_t, d.meshlet_info, ni, "meshletInfo");
  VG_FIELD(uint32_t, d.level_ranges, nr, "levelRanges");
  VG_FIELD(float, d.level_errors, ne, "levelErrors");
  VG_FIELD(float, d.level_guard_errors, ng, "levelGuardErrors");
  size_t nt = 0;
  VG_FIELD(uint32_t, d.meshlet_triangles, nt, "meshletTriangles");
  d.triangle_count = d.meshlet_triangles ? nt : 0;
  size_t npos = 0, nmv = 0;
  VG_FIELD(float, d.positions, npos, "positions");
  VG_FIELD(uint32_t, d.meshlet_vertices, nmv, "meshletVertices");
  d.vertex_count = d.positions ? static_cast<uint32_t>(npos / 3) : 0;
  d.meshlet_vertex_count = d.meshlet_vertices ? nmv : 0;
  d.meshlet_count = static_cast<uint32_t>(ni / 4);
  d.level_count = static_cast<uint32_t>(nr / 2);
  // Every array the library reads is as long as the counts taken from meshletInfo and levelRanges.
  if (ni % 4 || nr % 2 || npos % 3 || nb < static_cast<size_t>(d.meshlet_count) * 16 || ne < static_cast<size_t>(d.level_count) * 2 ||
      ng < static_cast<size_t>(d.level_count) * 2) {
    napi_throw_range_error(env, nullptr,
                           "vgcut: addMesh needs meshletBounds of 16, levelErrors and levelGuardErrors of 2 per level "
                           "(meshletInfo of 4 per meshlet, levelRanges of 2 per level, positions of 3 per vertex)");
    return nullptr;
  }
  const char* spheres[3] = {"boundingSphere", "lodBoundsSphere", "cullBoundsSphere"};
  float* dst[3] = {d.bounding_sphere, d.lod_bounds_sphere, d.cull_bounds_sphere};
  for (int k = 0; k < 3; k++) {
    const float* src = nullptr;
    VG_FIELD(float, src, n4, spheres[k]);
    if (src && n4 == 4) memcpy(dst[k], src, sizeof(float) * 4);
  }
#undef VG_FIELD
  uint32_t slot = 0;
  if (int rc = vg_mesh_add(s->scene, &d, &slot)) return throw_vg(env, rc, s->ctx);
  napi_value out;
  CHECK_NAPI(napi_create_uint32(env, slot, &out));
  return out;
}

// setInstances(scene, slot, matrices: Float32Array, colors?: Float32Array) -> Uint32Array gpuToInput
static napi_value SetInstances(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  size_t nm = 0, nc = 0;
  bool bad = false;
  const float* m = typed<float>(env, argv[2], &nm, &bad, "matrices");
  if (bad) return nullptr;
  const float* c = argc > 3 ? typed<float>(env, argv[3], &nc, &bad, "colors") : nullptr;
  if (bad) return nullptr;
  const uint32_t count = static_cast<uint32_t>(nm / 16);
  if (nm % 16 || (c && nc < static_cast<size_t>(count) * 4)) {
    napi_throw_range_error(env, nullptr, "vgcut: setInstances needs 16 floats per matrix and 4 per color");
    return nullptr;
  }
  napi_value ab, out;
  void* order = nullptr;
  CHECK_NAPI(napi_create_arraybuffer(env, sizeof(uint32_t) * count, &order, &ab));
  if (int rc = vg_instances_set(s->scene, slot, m, c, count, static_cast<uint32_t*>(order))) return throw_vg(env, rc, s->ctx);
  CHECK_NAPI(napi_create_typedarray(env, napi_uint32_array, count, ab, 0, &out));
  return out;
}

// setFlags(scene, slot, {maxDrawDistance, minPixelRadius, enabled, shadow, occlusion}) (GeometryPool.writeFlags)
static napi_value SetFlags(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  double max_draw = 1e30, min_px = 0.0;
  bool enabled = true, shadow = true, occlusion = false;
  napi_get_value_double(env, prop(env, argv[2], "maxDrawDistance"), &max_draw);
  napi_get_value_double(env, prop(env, argv[2], "minPixelRadius"), &min_px);
  napi_get_value_bool(env, prop(env, argv[2], "enabled"), &enabled);
  napi_get_value_bool(env, prop(env, argv[2], "shadow"), &shadow);
  napi_get_value_bool(env, prop(env, argv[2], "occlusion"), &occlusion);
  if (int rc = vg_mesh_flags(s->scene, slot, static_cast<float>(max_draw), static_cast<float>(min_px), enabled, shadow, occlusion))
    return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// removeMesh(scene, slot) (GeometryPool.remove, as VirtualGeometry.unregister calls it): the slot is freed
static napi_value RemoveMesh(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  if (int rc = vg_mesh_remove(s->scene, slot)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// setHzb(scene, viewSlot, depth: Float32Array, width, height): the slot's HZB from a depth image (row 0 at
// the top), as OcclusionCulling's copy and reduce nodes build it
static napi_value SetHzb(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0, width = 0, height = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  CHECK_NAPI(napi_get_value_uint32(env, argv[3], &width));
  CHECK_NAPI(napi_get_value_uint32(env, argv[4], &height));
  size_t n = 0;
  bool bad = false;
  const float* depth = typed<float>(env, argv[2], &n, &bad, "depth");
  if (bad) return nullptr;
  if (!depth || n < static_cast<size_t>(width) * height) {
    napi_throw_range_error(env, nullptr, "vgcut: depth array smaller than width x height");
    return nullptr;
  }
  if (int rc = vg_hzb_set(s->scene, slot, depth, width, height)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// The cut (and the second phase) run off the JS thread as napi_async_work and resolve a promise, so the
// render loop keeps going while the GPU cuts.
// The scene's external is referenced until the work completes (its finalizer cannot run mid-cut), and the
// context's in-flight flag is set for as long.
struct CutWork {
  Scene* s;
  std::vector<vg_view> views;
  bool phase2 = false;  // vg_cut_phase2 instead of the cut
  int rc = 0;
  napi_ref scene_ref = nullptr;
  napi_deferred deferred;
  napi_async_work work;
};

static napi_value QueueCut(napi_env env, napi_callback_info info, bool phase2) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* data = nullptr;
  size_t bytes = 0;
  uint32_t count = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_arraybuffer_info(env, argv[1], &data, &bytes));
  CHECK_NAPI(napi_get_value_uint32(env, argv[2], &count));
  if (bytes < sizeof(vg_view) * count) {
    napi_throw_range_error(env, nullptr, "vgcut: view buffer smaller than count records");
    return nullptr;
  }
  auto* w = new CutWork{s, std::vector<vg_view>(count)};
  w->phase2 = phase2;
  memcpy(w->views.data(), data, sizeof(vg_view) * count);
  if (napi_create_reference(env, argv[0], 1, &w->scene_ref) != napi_ok) {
    delete w;
    napi_throw_error(env, nullptr, "vgcut: N-API call failed");
    return nullptr;
  }
  napi_value promise, name;
  CHECK_NAPI(napi_create_promise(env, &w->deferred, &promise));
  CHECK_NAPI(napi_create_string_utf8(env, "vgcut.cut", NAPI_AUTO_LENGTH, &name));
  CHECK_NAPI(napi_create_async_work(
      env, nullptr, name,
      [](napi_env, void* d) {
        auto* w = static_cast<CutWork*>(d);
        const uint32_t n = static_cast<uint32_t>(w->views.size());
        w->rc = w->phase2 ? vg_cut_phase2(w->s->scene, w->views.data(), n) : vg_cut_graph(w->s->scene, w->views.data(), n);
        if (!w->rc) w->rc = vg_sync(w->s->scene);
      },
      [](napi_env env, napi_status, void* d) {
        auto* w = static_cast<CutWork*>(d);
        w->s->c->in_flight = false;
        napi_delete_reference(env, w->scene_ref);
        napi_value result;
        if (w->rc) {
          napi_value msg;
          napi_create_string_utf8(env, vg_last_error(w->s->ctx), NAPI_AUTO_LENGTH, &msg);
          napi_create_error(env, nullptr, msg, &result);
          napi_reject_deferred(env, w->deferred, result);
        } else {
          napi_get_undefined(env, &result);
          napi_resolve_deferred(env, w->deferred, result);
        }
        napi_delete_async_work(env, w->work);
        delete w;
      },
      w, &w->work));
  CHECK_NAPI(napi_queue_async_work(env, w->work));
  s->c->in_flight = true;
  return promise;
}

// cut(scene, views: ArrayBuffer of packed vg_view records, count) -> Promise (VirtualGeometry.update's compute)
static napi_value Cut(napi_env env, napi_callback_info info) { return QueueCut(env, info, false); }

// cutPhase2(scene, views, count) -> Promise: the second phase of two-phase occlusion (views with occlusion 2,
// after setting each slot's new HZB)
static napi_value CutPhase2(napi_env env, napi_callback_info info) { return QueueCut(env, info, true); }

// readPhase2(scene, viewSlot, counters: Uint32Array, drawArgs: Uint32Array): the second phase's counters and args
static napi_value ReadPhase2(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  size_t nc = 0, na = 0;
  bool bad = false;
  uint32_t* counters = typed<uint32_t>(env, argv[2], &nc, &bad, "counters");
  if (bad) return nullptr;
  uint32_t* args = typed<uint32_t>(env, argv[3], &na, &bad, "drawArgs");
  if (bad) return nullptr;
  if (!counters || !args || nc < vg_counter_words(s->ctx) || na < vg_draw_args_words(s->ctx)) {
    napi_throw_range_error(env, nullptr, "vgcut: counters or drawArgs array too small");
    return nullptr;
  }
  if (int rc = vg_read_phase2(s->scene, slot, counters, args)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// read(scene, viewSlot, counters: Uint32Array, drawArgs: Uint32Array, drawList: Uint32Array) (readCounters)
static napi_value Read(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  size_t nc = 0, na = 0, nl = 0;
  bool bad = false;
  uint32_t* counters = typed<uint32_t>(env, argv[2], &nc, &bad, "counters");
  if (bad) return nullptr;
  uint32_t* args = typed<uint32_t>(env, argv[3], &na, &bad, "drawArgs");
  if (bad) return nullptr;
  uint32_t* list = argc > 4 ? typed<uint32_t>(env, argv[4], &nl, &bad, "drawList") : nullptr;
  if (bad) return nullptr;
  if (!counters || !args || nc < vg_counter_words(s->ctx) || na < vg_draw_args_words(s->ctx)) {
    napi_throw_range_error(env, nullptr, "vgcut: counters or drawArgs array too small");
    return nullptr;
  }
  if (int rc = vg_read(s->scene, slot, counters, args, list, static_cast<uint32_t>(nl / 8), nullptr)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// readIndices(scene, slot, cut, first, out: Uint32Array): the expand pass's index buffer of a cut
// (cameraIndex / shadowIndex), out.length indices from `first`
static napi_value ReadIndices(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t slot = 0, cut = 0;
  double first = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  CHECK_NAPI(napi_get_value_uint32(env, argv[2], &cut));
  CHECK_NAPI(napi_get_value_double(env, argv[3], &first));
  size_t n = 0;
  bool bad = false;
  uint32_t* out = typed<uint32_t>(env, argv[4], &n, &bad, "out");
  if (bad) return nullptr;
  if (!out || first < 0) {
    napi_throw_type_error(env, nullptr, "vgcut: readIndices needs a Uint32Array and first >= 0");
    return nullptr;
  }
  if (int rc = vg_index_read(s->scene, slot, cut, static_cast<uint64_t>(first), n, out)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// poolRead(scene, buffer) -> ArrayBuffer copy of a pool buffer in the reference layout (VG_POOL_*)
static napi_value PoolRead(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  uint32_t which = 0;
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &which));
  uint64_t bytes = 0;
  if (int rc = vg_scene_pool_read(s->scene, which, nullptr, 0, &bytes)) return throw_vg(env, rc, s->ctx);
  napi_value ab;
  void* data = nullptr;
  CHECK_NAPI(napi_create_arraybuffer(env, bytes, &data, &ab));
  if (int rc = vg_scene_pool_read(s->scene, which, data, bytes, &bytes)) return throw_vg(env, rc, s->ctx);
  return ab;
}

// counterWords(ctx), drawArgsWords(ctx): array sizes for read()
static napi_value CounterWords(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  CHECK_NAPI(napi_create_uint32(env, vg_counter_words(c->ctx), &out));
  return out;
}

static napi_value DrawArgsWords(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Ctx* c = ctx_of(env, argv[0]);
  if (!c) return nullptr;
  CHECK_NAPI(napi_create_uint32(env, vg_draw_args_words(c->ctx), &out));
  return out;
}

// hzbFromCut(scene, slot, viewProj: Float32Array(16), width, height, withPhase2?) -- the occluder pass of
// OcclusionCulling.beginFrame (OcclusionCulling.ts:244-273) on the device, queued on the scene's stream
static napi_value HzbFromCut(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  uint32_t slot = 0, width = 0, height = 0;
  bool with_phase2 = false;
  CHECK_NAPI(napi_get_value_uint32(env, argv[1], &slot));
  size_t n = 0;
  bool bad = false;
  const float* vp = typed<float>(env, argv[2], &n, &bad, "viewProj");
  if (bad) return nullptr;
  if (!vp || n != 16) {
    napi_throw_type_error(env, nullptr, "hzbFromCut: viewProj must be a Float32Array of 16");
    return nullptr;
  }
  CHECK_NAPI(napi_get_value_uint32(env, argv[3], &width));
  CHECK_NAPI(napi_get_value_uint32(env, argv[4], &height));
  if (argc > 5) napi_get_value_bool(env, argv[5], &with_phase2);
  if (int rc = vg_hzb_from_cut(s->scene, slot, vp, width, height, with_phase2 ? 1u : 0u)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// hzbFromCuts(scene, rasters: [{slot, viewProj: Float32Array(16), width, height, withPhase2?}]) -- hzbFromCut for
// several view slots in one launch sequence (vg_hzb_from_cuts)
static napi_value HzbFromCuts(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Scene* s = scene_of(env, argv[0]);
  if (!s) return nullptr;
  bool is_array = false;
  if (argc < 2 || napi_is_array(env, argv[1], &is_array) != napi_ok || !is_array) {
    napi_throw_type_error(env, nullptr, "hzbFromCuts: rasters must be an array");
    return nullptr;
  }
  uint32_t count = 0;
  CHECK_NAPI(napi_get_array_length(env, argv[1], &count));
  std::vector<vg_raster> rasters(count);
  for (uint32_t i = 0; i < count; i++) {
    napi_value item, v;
    CHECK_NAPI(napi_get_element(env, argv[1], i, &item));
    vg_raster& q = rasters[i];
    std::memset(&q, 0, sizeof(q));
    CHECK_NAPI(napi_get_named_property(env, item, "slot", &v));
    CHECK_NAPI(napi_get_value_uint32(env, v, &q.view_slot));
    CHECK_NAPI(napi_get_named_property(env, item, "width", &v));
    CHECK_NAPI(napi_get_value_uint32(env, v, &q.width));
    CHECK_NAPI(napi_get_named_property(env, item, "height", &v));
    CHECK_NAPI(napi_get_value_uint32(env, v, &q.height));
    bool has_phase2 = false, with_phase2 = false;
    if (napi_has_named_property(env, item, "withPhase2", &has_phase2) == napi_ok && has_phase2 &&
        napi_get_named_property(env, item, "withPhase2", &v) == napi_ok)
      napi_get_value_bool(env, v, &with_phase2);
    q.with_phase2 = with_phase2 ? 1u : 0u;
    CHECK_NAPI(napi_get_named_property(env, item, "viewProj", &v));
    size_t n = 0;
    bool bad = false;
    const float* vp = typed<float>(env, v, &n, &bad, "viewProj");
    if (bad) return nullptr;
    if (!vp || n != 16) {
      napi_throw_type_error(env, nullptr, "hzbFromCuts: viewProj must be a Float32Array of 16");
      return nullptr;
    }
    std::memcpy(q.view_proj, vp, sizeof(q.view_proj));
  }
  if (int rc = vg_hzb_from_cuts(s->scene, rasters.data(), count)) return throw_vg(env, rc, s->ctx);
  return nullptr;
}

// A new typed array holding a copy of `n` elements at `src`.
static napi_value copy_typed(napi_env env, napi_typedarray_type type, const void* src, size_t n, size_t elem) {
  void* data = nullptr;
  napi_value ab, out;
  if (napi_create_arraybuffer(env, n * elem, &data, &ab) != napi_ok) return nullptr;
  if (n) memcpy(data, src, n * elem);
  if (napi_create_typedarray(env, type, n, ab, 0, &out) != napi_ok) return nullptr;
  return out;
}

// buildVirtualMesh(src, options?) -> VirtualMeshData, src = {positions, indices, normals?, colors?, uvs?}
// (buildVirtualMesh.ts:185 with VirtualMeshSource), or buildVirtualMesh(positions, indices, options?);
// options take the reference's names.
static napi_value BuildVirtualMesh(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) {
    napi_throw_type_error(env, nullptr, "buildVirtualMesh: expected a source");
    return nullptr;
  }
  bool is_typed = false;
  CHECK_NAPI(napi_is_typedarray(env, argv[0], &is_typed));
  napi_value src_obj = nullptr, options = nullptr;
  napi_value pos_v = argv[0], idx_v = argc > 1 ? argv[1] : nullptr;
  if (!is_typed) {
    napi_valuetype t;
    CHECK_NAPI(napi_typeof(env, argv[0], &t));
    if (t != napi_object) {
      napi_throw_type_error(env, nullptr, "buildVirtualMesh: expected a source object or typed arrays");
      return nullptr;
    }
    src_obj = argv[0];
    pos_v = prop(env, src_obj, "positions");
    idx_v = prop(env, src_obj, "indices");
    if (argc > 1) options = argv[1];
  } else if (argc > 2) {
    options = argv[2];
  }
  if (!idx_v) {
    napi_throw_type_error(env, nullptr, "buildVirtualMesh: indices missing");
    return nullptr;
  }
  size_t np = 0, ni = 0;
  bool bad = false;
  const float* positions = typed<float>(env, pos_v, &np, &bad, "positions");
  if (bad) return nullptr;
  const uint32_t* indices = typed<uint32_t>(env, idx_v, &ni, &bad, "indices");
  if (bad) return nullptr;
  if (!positions || !indices || np % 3 || ni % 3) {
    napi_throw_type_error(env, nullptr, "buildVirtualMesh: positions must be a Float32Array of 3 per vertex, indices a Uint32Array of 3 per triangle");
    return nullptr;
  }
  const size_t nv = np / 3;
  const float* attr[3] = {nullptr, nullptr, nullptr};
  if (src_obj) {
    const char* names[3] = {"normals", "colors", "uvs"};
    const size_t per[3] = {3, 3, 2};
    for (int a = 0; a < 3; a++) {
      napi_value v = prop(env, src_obj, names[a]);
      napi_valuetype t;
      if (!v || napi_typeof(env, v, &t) != napi_ok || t == napi_undefined || t == napi_null) continue;
      size_t n = 0;
      attr[a] = typed<float>(env, v, &n, &bad, names[a]);
      if (bad) return nullptr;
      if (!attr[a] || n != nv * per[a]) {
        napi_throw_range_error(env, nullptr, "buildVirtualMesh: an attribute does not cover every vertex");
        return nullptr;
      }
    }
  }
  vgd_options o;
  vgd_default_options(&o);
  if (options) {
    napi_valuetype t;
    CHECK_NAPI(napi_typeof(env, options, &t));
    if (t == napi_object) {
      auto u32 = [&](const char* name, uint32_t* out) {
        uint32_t x;
        if (napi_get_value_uint32(env, prop(env, options, name), &x) == napi_ok) *out = x;
      };
      auto f64 = [&](const char* name, double* out) {
        double x;
        if (napi_get_value_double(env, prop(env, options, name), &x) == napi_ok) *out = x;
      };
      auto flag = [&](const char* name, int32_t* out) {
        bool x;
        if (napi_get_value_bool(env, prop(env, options, name), &x) == napi_ok) *out = x ? 1 : 0;
      };
      u32("groupSize", &o.group_size);
      u32("maxLodLevels", &o.max_lod_levels);
      f64("maxGroupKeepRatio", &o.max_group_keep_ratio);
      f64("simplifyRatio", &o.simplify_ratio);
      f64("minLevelReduction", &o.min_level_reduction);
      flag("prune", &o.prune);
      flag("tailSimplify", &o.tail_simplify);
      u32("minRootTriangles", &o.min_root_triangles);
      u32("threads", &o.threads);
      flag("voxelLods", &o.voxel_lods);
      u32("voxelResolution", &o.voxel_resolution);
      flag("aggregateLods", &o.aggregate_lods);
    }
  }
  vgd_source src{static_cast<uint32_t>(nv), positions, attr[0], attr[1], attr[2], ni, indices};
  vgd_mesh* mesh = nullptr;
  if (int rc = vgd_build_source(&src, &o, &mesh)) {
    napi_throw_error(env, nullptr, rc == -1 ? "buildVirtualMesh: invalid mesh" : "buildVirtualMesh: build failed");
    return nullptr;
  }
  vgd_mesh_view v;
  vgd_mesh_get(mesh, &v);
  napi_value out, stats, x;
  CHECK_NAPI(napi_create_object(env, &out));
  auto put = [&](napi_value obj, const char* name, napi_value val) { napi_set_named_property(env, obj, name, val); };
  const size_t n = v.meshlet_count;
  put(out, "positions", copy_typed(env, napi_float32_array, v.positions, static_cast<size_t>(v.vertex_count) * 3, 4));
  napi_value null_value;
  napi_get_null(env, &null_value);
  put(out, "normals", v.normals ? copy_typed(env, napi_float32_array, v.normals, static_cast<size_t>(v.vertex_count) * 3, 4) : null_value);
  put(out, "colors", v.colors ? copy_typed(env, napi_float32_array, v.colors, static_cast<size_t>(v.vertex_count) * 3, 4) : null_value);
  put(out, "uvs", v.uvs ? copy_typed(env, napi_float32_array, v.uvs, static_cast<size_t>(v.vertex_count) * 2, 4) : null_value);
  put(out, "indices", copy_typed(env, napi_uint32_array, v.indices, v.index_count, 4));
  napi_create_uint32(env, v.meshlet_count, &x);
  put(out, "meshletCount", x);
  put(out, "meshletBounds", copy_typed(env, napi_float32_array, v.meshlet_bounds, n * 16, 4));
  put(out, "meshletInfo", copy_typed(env, napi_uint32_array, v.meshlet_info, n * 4, 4));
  put(out, "meshletVertices", copy_typed(env, napi_uint32_array, v.meshlet_vertices, v.meshlet_vertex_count, 4));
  put(out, "meshletTriangles", copy_typed(env, napi_uint32_array, v.meshlet_triangles, v.index_count / 3, 4));
  put(out, "levelRanges", copy_typed(env, napi_uint32_array, v.level_ranges, v.level_count * 2, 4));
  put(out, "levelErrors", copy_typed(env, napi_float32_array, v.level_errors, v.level_count * 2, 4));
  put(out, "levelGuardErrors", copy_typed(env, napi_float32_array, v.level_guard_errors, v.level_count * 2, 4));
  put(out, "replacementGroup", copy_typed(env, napi_int32_array, v.replacement_group, n, 4));
  put(out, "replacementStart", copy_typed(env, napi_uint32_array, v.replacement_start, v.group_count + 1, 4));
  put(out, "replacementIndices",
      copy_typed(env, napi_uint32_array, v.replacement_indices, v.replacement_start[v.group_count], 4));
  put(out, "boundingSphere", copy_typed(env, napi_float32_array, v.bounding_sphere, 4, 4));
  put(out, "lodBoundsSphere", copy_typed(env, napi_float32_array, v.lod_bounds_sphere, 4, 4));
  put(out, "cullBoundsSphere", copy_typed(env, napi_float32_array, v.cull_bounds_sphere, 4, 4));
  CHECK_NAPI(napi_create_object(env, &stats));
  napi_create_double(env, static_cast<double>(v.leaf_triangles), &x);
  put(stats, "leafTriangles", x);
  napi_create_uint32(env, v.leaf_meshlets, &x);
  put(stats, "leafMeshlets", x);
  napi_create_uint32(env, v.root_triangles, &x);
  put(stats, "rootTriangles", x);
  napi_create_uint32(env, v.root_meshlets, &x);
  put(stats, "rootMeshlets", x);
  napi_create_uint32(env, v.lod_levels, &x);
  put(stats, "lodLevels", x);
  napi_create_uint32(env, v.voxel_levels, &x);
  put(stats, "voxelLevels", x);
  napi_create_uint32(env, v.aggregate_groups, &x);
  put(stats, "aggregateGroups", x);
  put(out, "stats", stats);
  vgd_mesh_free(mesh);
  return out;
}

static napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"createContext", nullptr, CreateContext, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"createScene", nullptr, CreateScene, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"addMesh", nullptr, AddMesh, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setInstances", nullptr, SetInstances, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setFlags", nullptr, SetFlags, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"removeMesh", nullptr, RemoveMesh, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cut", nullptr, Cut, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"read", nullptr, Read, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"poolRead", nullptr, PoolRead, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readIndices", nullptr, ReadIndices, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"cutPhase2", nullptr, CutPhase2, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"setHzb", nullptr, SetHzb, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"readPhase2", nullptr, ReadPhase2, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"counterWords", nullptr, CounterWords, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"drawArgsWords", nullptr, DrawArgsWords, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"buildVirtualMesh", nullptr, BuildVirtualMesh, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"hzbFromCut", nullptr, HzbFromCut, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
      {"hzbFromCuts", nullptr, HzbFromCuts, nullptr, nullptr, nullptr, napi_enumerable, nullptr},
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
