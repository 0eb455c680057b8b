"""ctypes bindings of the C ABI in include/vgcut.h (lib/libvgcut.so).

The library is required: importing the bindings without it raises, so no caller can silently run
without the HIP path.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import _native

VG_OK = 0
VG_E_INVALID = -1
VG_E_NOMEM = -2
VG_E_CAPACITY = -3
VG_E_DEVICE = -4
VG_SERVE_SHORTEST = 0
VG_SERVE_FIFO = 1

VG_POOL_MESH_TABLE = 0
VG_POOL_INSTANCES = 1
VG_POOL_CELLS = 2
VG_POOL_MESHLET_BOUNDS = 3
VG_POOL_MESHLET_INFO = 4
VG_POOL_MESHLET_TRIANGLES = 5
VG_POOL_VERTICES = 6
VG_POOL_MESHLET_VERTICES = 7

LIBRARY = _native.library_path("libvgcut.so")

# Every entry point include/vgcut.h declares (the ABI test checks the library exports them).
EXPORTS = (
    "vg_create", "vg_destroy", "vg_last_error", "vg_scene_create", "vg_scene_destroy", "vg_scene_bytes", "vg_mesh_add",
    "vg_instances_set", "vg_mesh_flags", "vg_mesh_remove", "vg_hzb_set", "vg_hzb_from_cut", "vg_hzb_from_cuts", "vg_hzb_invalidate", "vg_hzb_read", "vg_cut", "vg_cut_graph", "vg_sync", "vg_read",
    "vg_state_get", "vg_state_set", "vg_scene_pool_read", "vg_scene_mesh_slots", "vg_mesh_instance_base",
    "vg_counter_words", "vg_draw_args_words", "vg_profile", "vg_profile_read", "vg_footprint", "vg_budget_thresholds",
    "vg_schedule_get", "vg_schedule_set", "vg_tune", "vg_index_read", "vg_index_buffer",
    "vg_cut_phase2", "vg_read_phase2", "vg_phase_rejects", "vg_cut_static",
    "vg_server_create", "vg_server_destroy", "vg_server_last_error", "vg_server_add_scene", "vg_server_submit", "vg_server_poll",
    "vg_server_read", "vg_server_release", "vg_server_stats_get",
)


class VgError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"vgcut error {code}: {message}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("max_triangles", ctypes.c_uint32),
        ("max_shadow_triangles", ctypes.c_uint32),
        ("max_meshlets", ctypes.c_uint32),
        ("max_views", ctypes.c_uint32),
        ("mesh_slots", ctypes.c_uint32),
        ("hzb_max_size", ctypes.c_uint32),
        ("budget_bins", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("expand", ctypes.c_uint32),
    ]


class MeshDesc(ctypes.Structure):
    _fields_ = [
        ("meshlet_count", ctypes.c_uint32),
        ("meshlet_bounds", ctypes.c_void_p),
        ("meshlet_info", ctypes.c_void_p),
        ("level_count", ctypes.c_uint32),
        ("level_ranges", ctypes.c_void_p),
        ("level_errors", ctypes.c_void_p),
        ("level_guard_errors", ctypes.c_void_p),
        ("bounding_sphere", ctypes.c_float * 4),
        ("lod_bounds_sphere", ctypes.c_float * 4),
        ("cull_bounds_sphere", ctypes.c_float * 4),
        ("meshlet_triangles", ctypes.c_void_p),
        ("triangle_count", ctypes.c_uint64),
        ("positions", ctypes.c_void_p),
        ("vertex_count", ctypes.c_uint32),
        ("meshlet_vertices", ctypes.c_void_p),
        ("meshlet_vertex_count", ctypes.c_uint64),
    ]


class View(ctypes.Structure):
    _fields_ = [
        ("view_matrix", ctypes.c_float * 16),
        ("projection_matrix", ctypes.c_float * 16),
        ("frustum_planes", ctypes.c_float * 24),
        ("shadow_sweep", ctypes.c_float * 3),
        ("proj_scale", ctypes.c_float),
        ("near_plane", ctypes.c_float),
        ("error_threshold", ctypes.c_float),
        ("shadow_error_scale", ctypes.c_float),
        ("lod_blend_max", ctypes.c_float),
        ("lod_scale_limit", ctypes.c_float),
        ("min_pixel_radius", ctypes.c_float),
        ("frustum_culling", ctypes.c_uint32),
        ("occlusion", ctypes.c_uint32),
        ("slot", ctypes.c_uint32),
        ("triangle_budget", ctypes.c_uint32),
        ("threshold_range", ctypes.c_float * 2),
        ("max_error_threshold", ctypes.c_float),
    ]


class Raster(ctypes.Structure):
    """vg_raster: one raster of vg_hzb_from_cuts."""
    _fields_ = [("view_slot", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("with_phase2", ctypes.c_uint32), ("view_proj", ctypes.c_float * 16)]


class Schedule(ctypes.Structure):
    """Launch schedule of the persistent passes (vg_schedule): grids per CU and the meshlet-pass variant."""

    _fields_ = [
        ("instance_blocks_per_cu", ctypes.c_uint32),
        ("meshlet_blocks_per_cu", ctypes.c_uint32),
        ("meshlet_waves_per_simd", ctypes.c_uint32),
    ]

    def as_tuple(self):
        return (self.instance_blocks_per_cu, self.meshlet_blocks_per_cu, self.meshlet_waves_per_simd)


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_uint64),
        ("lanes", ctypes.c_uint32),
        ("max_batch_views", ctypes.c_uint32),
        ("max_batch_requests", ctypes.c_uint32),
        ("max_bypass_us", ctypes.c_uint32),
        ("first_work_items", ctypes.c_uint32),
        ("policy", ctypes.c_uint32),
        ("max_blocks_per_cu", ctypes.c_uint32),
    ]


class RequestDone(ctypes.Structure):
    _fields_ = [
        ("ticket", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("views", ctypes.c_uint32),
        ("batch_requests", ctypes.c_uint32),
        ("lane", ctypes.c_uint32),
        ("queued_us", ctypes.c_double),
        ("run_us", ctypes.c_double),
        ("total_us", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_uint64),
        ("resident_bytes", ctypes.c_uint64),
        ("inflight_bytes", ctypes.c_uint64),
        ("peak_inflight_bytes", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("requests_done", ctypes.c_uint64),
        ("max_batch_requests_seen", ctypes.c_uint64),
        ("reruns", ctypes.c_uint64),
        ("queued", ctypes.c_uint32),
        ("running", ctypes.c_uint32),
    ]


class Metrics(ctypes.Structure):
    _fields_ = [
        ("cells_tested", ctypes.c_uint64),
        ("instances_tested", ctypes.c_uint64),
        ("work_items", ctypes.c_uint64 * 2),
        ("clusters_tested", ctypes.c_uint64 * 2),
        ("occlusion_tests", ctypes.c_uint64),
        ("error_threshold", ctypes.c_float),
        ("budget_bin", ctypes.c_uint32),
        ("budget_cluster_tests", ctypes.c_uint64),
    ]


_lib = None


def load():
    """Loads lib/libvgcut.so (raises when it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBRARY):
        raise RuntimeError(f"{LIBRARY} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(LIBRARY)
    P, U32, I32, U64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    sig = {
        "vg_create": ([ctypes.POINTER(Config), ctypes.POINTER(P)], ctypes.c_int),
        "vg_destroy": ([P], None),
        "vg_last_error": ([P], ctypes.c_char_p),
        "vg_scene_create": ([P, ctypes.POINTER(P)], ctypes.c_int),
        "vg_scene_destroy": ([P], None),
        "vg_scene_bytes": ([P], U64),
        "vg_mesh_add": ([P, ctypes.POINTER(MeshDesc), ctypes.POINTER(U32)], ctypes.c_int),
        "vg_instances_set": ([P, U32, P, P, U32, P], ctypes.c_int),
        "vg_mesh_flags": ([P, U32, ctypes.c_float, ctypes.c_float, U32, U32, U32], ctypes.c_int),
        "vg_mesh_remove": ([P, U32], ctypes.c_int),
        "vg_hzb_set": ([P, U32, P, U32, U32], ctypes.c_int),
        "vg_hzb_from_cut": ([P, U32, P, U32, U32, U32], ctypes.c_int),
        "vg_hzb_from_cuts": ([P, P, U32], ctypes.c_int),
        "vg_hzb_invalidate": ([P, U32], ctypes.c_int),
        "vg_hzb_read": ([P, U32, P, U64, ctypes.POINTER(U64)], ctypes.c_int),
        "vg_cut": ([P, P, U32], ctypes.c_int),
        "vg_cut_graph": ([P, P, U32], ctypes.c_int),
        "vg_cut_static": ([P, P, U32], ctypes.c_int),
        "vg_sync": ([P], ctypes.c_int),
        "vg_read": ([P, U32, P, P, P, U32, ctypes.POINTER(Metrics)], ctypes.c_int),
        "vg_state_get": ([P, U32, P], ctypes.c_int),
        "vg_state_set": ([P, U32, P], ctypes.c_int),
        "vg_scene_pool_read": ([P, U32, P, U64, ctypes.POINTER(U64)], ctypes.c_int),
        "vg_scene_mesh_slots": ([P], U32),
        "vg_mesh_instance_base": ([P, U32, ctypes.POINTER(U32), ctypes.POINTER(U32)], ctypes.c_int),
        "vg_counter_words": ([P], U32),
        "vg_draw_args_words": ([P], U32),
        "vg_arith_probe": ([P, P, P, U32], ctypes.c_int),
        "vg_scan_probe": ([P, P, U32], ctypes.c_int),
        "vg_profile": ([P, U32], ctypes.c_int),
        "vg_budget_thresholds": ([ctypes.c_float, ctypes.c_float, U32, P], ctypes.c_int),
        "vg_profile_read": ([P, P, U32, ctypes.POINTER(U64), P], ctypes.c_int),
        "vg_footprint": ([P, P, U32, ctypes.POINTER(U64), ctypes.POINTER(U64)], ctypes.c_int),
        "vg_server_create": ([P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)], ctypes.c_int),
        "vg_server_destroy": ([P], None),
        "vg_server_last_error": ([P], ctypes.c_char_p),
        "vg_server_add_scene": ([P, P, ctypes.POINTER(U32)], ctypes.c_int),
        "vg_server_submit": ([P, U32, P, U32, ctypes.POINTER(U64)], ctypes.c_int),
        "vg_server_poll": ([P, ctypes.POINTER(RequestDone), U32, ctypes.POINTER(U32), U32], ctypes.c_int),
        "vg_server_read": ([P, U64, U32, P, P, P, U32, ctypes.POINTER(Metrics)], ctypes.c_int),
        "vg_server_release": ([P, U64], ctypes.c_int),
        "vg_server_stats_get": ([P, ctypes.POINTER(ServerStats)], ctypes.c_int),
        "vg_schedule_get": ([P, ctypes.POINTER(Schedule)], ctypes.c_int),
        "vg_schedule_set": ([P, ctypes.POINTER(Schedule)], ctypes.c_int),
        "vg_tune": ([P, P, U32, U32, ctypes.POINTER(Schedule), ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
        "vg_index_read": ([P, U32, U32, U64, U64, P], ctypes.c_int),
        "vg_cut_phase2": ([P, P, U32], ctypes.c_int),
        "vg_read_phase2": ([P, U32, P, P], ctypes.c_int),
        "vg_phase_rejects": ([P, U32, ctypes.POINTER(U32), ctypes.POINTER(U32)], ctypes.c_int),
        "vg_index_buffer": ([P, U32, U32, ctypes.POINTER(P), ctypes.POINTER(U64)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    del I32
    _lib = lib
    return lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class Context:
    """A vg_ctx: capacities and the HIP stream (VirtualGeometry constructor options)."""

    def __init__(self, max_triangles=10_000_000, max_shadow_triangles=4_000_000, max_meshlets=1 << 20, max_views=1,
                 mesh_slots=64, hzb_max_size=1024, budget_bins=256, device=0, expand=False):
        self.lib = load()
        cfg = Config(max_triangles, max_shadow_triangles, max_meshlets, max_views, mesh_slots, hzb_max_size, budget_bins, device,
                     1 if expand else 0)
        self.config = cfg
        h = ctypes.c_void_p()
        rc = self.lib.vg_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != VG_OK:
            raise VgError(rc, "vg_create failed (no HIP device?)")
        self.handle = h

    def check(self, rc: int, what: str = ""):
        if rc != VG_OK:
            msg = self.lib.vg_last_error(self.handle)
            raise VgError(rc, f"{what}: {msg.decode() if msg else ''}")

    @property
    def counter_words(self) -> int:
        return int(self.lib.vg_counter_words(self.handle))

    @property
    def draw_args_words(self) -> int:
        return int(self.lib.vg_draw_args_words(self.handle))

    def close(self):
        if self.handle:
            self.lib.vg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A vg_scene: one resident GeometryPool."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        ctx.check(self.lib.vg_scene_create(ctx.handle, ctypes.byref(h)), "vg_scene_create")
        self.handle = h

    def add_mesh(self, data) -> int:
        bounds = np.ascontiguousarray(data.meshletBounds, dtype=np.float32)
        info = np.ascontiguousarray(data.meshletInfo, dtype=np.uint32)
        ranges = np.ascontiguousarray(data.levelRanges, dtype=np.uint32)
        errors = np.ascontiguousarray(data.levelErrors, dtype=np.float32)
        guard = np.ascontiguousarray(data.levelGuardErrors, dtype=np.float32)
        d = MeshDesc()
        d.meshlet_count = data.meshletCount
        d.meshlet_bounds = bounds.ctypes.data
        d.meshlet_info = info.ctypes.data
        d.level_count = len(ranges) // 2
        d.level_ranges = ranges.ctypes.data
        d.level_errors = errors.ctypes.data
        d.level_guard_errors = guard.ctypes.data
        for i in range(4):
            d.bounding_sphere[i] = float(data.boundingSphere[i])
            d.lod_bounds_sphere[i] = float(data.lodBoundsSphere[i])
            d.cull_bounds_sphere[i] = float(data.cullBoundsSphere[i])
        tri = getattr(data, "meshletTriangles", None)
        if tri is not None:
            tri = np.ascontiguousarray(tri, dtype=np.uint32)
            d.meshlet_triangles = tri.ctypes.data
            d.triangle_count = len(tri)
        pos = getattr(data, "positions", None)
        mverts = getattr(data, "meshletVertices", None)
        if pos is not None and mverts is not None and tri is not None:
            pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
            mverts = np.ascontiguousarray(mverts, dtype=np.uint32)
            d.positions = pos.ctypes.data
            d.vertex_count = len(pos) // 3
            d.meshlet_vertices = mverts.ctypes.data
            d.meshlet_vertex_count = len(mverts)
        slot = ctypes.c_uint32()
        self.ctx.check(self.lib.vg_mesh_add(self.handle, ctypes.byref(d), ctypes.byref(slot)), "vg_mesh_add")
        return int(slot.value)

    def set_instances(self, slot: int, matrices: np.ndarray, colors: Optional[np.ndarray] = None) -> np.ndarray:
        m = np.ascontiguousarray(matrices, dtype=np.float32).reshape(-1, 16)
        c = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 4)
        order = np.empty(len(m), dtype=np.uint32)
        self.ctx.check(self.lib.vg_instances_set(self.handle, slot, _ptr(m), _ptr(c), len(m), _ptr(order)), "vg_instances_set")
        return order

    def remove_mesh(self, slot: int):
        """GeometryPool.remove (vg_mesh_remove): frees the slot; its instances and cells are dropped."""
        self.ctx.check(self.lib.vg_mesh_remove(self.handle, slot), "vg_mesh_remove")

    def set_flags(self, slot: int, max_draw_distance=1e30, min_pixel_radius=0.0, enabled=True, shadow=True, occlusion=False):
        self.ctx.check(self.lib.vg_mesh_flags(self.handle, slot, float(max_draw_distance), float(min_pixel_radius), int(bool(enabled)),
                                              int(bool(shadow)), int(bool(occlusion))), "vg_mesh_flags")

    def set_hzb(self, view_slot: int, depth: np.ndarray):
        d = np.ascontiguousarray(depth, dtype=np.float32)
        h, w = d.shape
        self.ctx.check(self.lib.vg_hzb_set(self.handle, view_slot, d.ctypes.data, w, h), "vg_hzb_set")

    def hzb_from_cut(self, view_slot: int, view_proj, width: int, height: int, with_phase2: bool = False):
        """The occluder pass on the device: the slot's last camera cut drawn depth-only with `view_proj`
        (column-major) into a width x height depth image, and the slot's HZB built from it."""
        vp = np.ascontiguousarray(np.asarray(view_proj, dtype=np.float32).reshape(16))
        self.ctx.check(self.lib.vg_hzb_from_cut(self.handle, view_slot, vp.ctypes.data, width, height, int(bool(with_phase2))),
                       "vg_hzb_from_cut")

    def hzb_from_cuts(self, rasters):
        """vg_hzb_from_cuts: [(view_slot, view_proj, width, height, with_phase2)] drawn in one launch sequence."""
        arr = (Raster * len(rasters))()
        for i, (slot, vp, w, h, phase2) in enumerate(rasters):
            arr[i].view_slot, arr[i].width, arr[i].height, arr[i].with_phase2 = slot, w, h, int(bool(phase2))
            for k, x in enumerate(np.asarray(vp, dtype=np.float32).reshape(16)):
                arr[i].view_proj[k] = float(x)
        self.ctx.check(self.lib.vg_hzb_from_cuts(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(rasters)), "vg_hzb_from_cuts")

    def hzb(self, view_slot: int) -> np.ndarray:
        n = ctypes.c_uint64()
        self.ctx.check(self.lib.vg_hzb_read(self.handle, view_slot, None, 0, ctypes.byref(n)), "vg_hzb_read")
        out = np.empty(n.value, dtype=np.float32)
        if n.value:
            self.ctx.check(self.lib.vg_hzb_read(self.handle, view_slot, out.ctypes.data, n.value, None), "vg_hzb_read")
        return out

    def invalidate_hzb(self, view_slot: int):
        self.ctx.check(self.lib.vg_hzb_invalidate(self.handle, view_slot), "vg_hzb_invalidate")

    def cut(self, views, graph: bool = False):
        arr = (View * len(views))(*views)
        fn = self.lib.vg_cut_graph if graph else self.lib.vg_cut
        self.ctx.check(fn(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(views)), "vg_cut")
        self._keep = arr

    def cut_static(self, views):
        """vg_cut_static: the cut with its launch buffers in the library's static scratch (no allocation;
        waits for the cut)."""
        arr = (View * len(views))(*views)
        self.ctx.check(self.lib.vg_cut_static(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(views)), "vg_cut_static")
        self._keep = arr

    def sync(self):
        self.ctx.check(self.lib.vg_sync(self.handle), "vg_sync")

    def cut_phase2(self, views):
        """Second phase of two-phase occlusion for the views of the last cut (occlusion = 2), against the
        slots' current HZBs."""
        arr = (View * len(views))(*views)
        self.ctx.check(self.lib.vg_cut_phase2(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(views)), "vg_cut_phase2")
        self._keep = arr

    def read_phase2(self, view_slot: int):
        """The second phase's counters and draw args of a view slot."""
        counters = np.empty(self.ctx.counter_words, dtype=np.uint32)
        args = np.empty(self.ctx.draw_args_words, dtype=np.uint32)
        self.ctx.check(self.lib.vg_read_phase2(self.handle, view_slot, counters.ctypes.data, args.ctypes.data), "vg_read_phase2")
        return counters, args

    def phase_rejects(self, view_slot: int):
        """(meshlet rejects recorded by the slot's last cut, whether more were dropped)."""
        n, dropped = ctypes.c_uint32(), ctypes.c_uint32()
        self.ctx.check(self.lib.vg_phase_rejects(self.handle, view_slot, ctypes.byref(n), ctypes.byref(dropped)), "vg_phase_rejects")
        return int(n.value), bool(dropped.value)

    def read(self, view_slot: int, max_entries: Optional[int] = None):
        """Counters, draw args, draw lists (camera, shadow) and metrics of a view slot."""
        if max_entries is None:
            max_entries = self.ctx.config.max_meshlets
        counters = np.empty(self.ctx.counter_words, dtype=np.uint32)
        args = np.empty(self.ctx.draw_args_words, dtype=np.uint32)
        draws = np.empty((2, max_entries, 4), dtype=np.uint32)
        met = Metrics()
        self.ctx.check(self.lib.vg_read(self.handle, view_slot, counters.ctypes.data, args.ctypes.data, draws.ctypes.data, max_entries,
                                        ctypes.byref(met)), "vg_read")
        return counters, args, draws, met

    def index_capacity(self, cut: int) -> int:
        """Capacity (u32 indices) of a cut's index buffer: triangles x 3 (contexts with the expand pass)."""
        cap = ctypes.c_uint64()
        self.ctx.check(self.lib.vg_index_buffer(self.handle, 0, cut, None, ctypes.byref(cap)), "vg_index_buffer")
        return int(cap.value)

    def read_indices(self, view_slot: int, cut: int, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        """`count` indices from `first` of a view slot's index buffer for `cut` (all of it by default)."""
        if count is None:
            count = self.index_capacity(cut) - first
        out = np.empty(count, dtype=np.uint32)
        self.ctx.check(self.lib.vg_index_read(self.handle, view_slot, cut, first, count, out.ctypes.data), "vg_index_read")
        return out

    def state(self, view_slot: int) -> np.ndarray:
        s = np.empty(8, dtype=np.uint32)
        self.ctx.check(self.lib.vg_state_get(self.handle, view_slot, s.ctypes.data), "vg_state_get")
        return s

    def set_state(self, view_slot: int, state: np.ndarray):
        s = np.ascontiguousarray(state, dtype=np.uint32)
        self.ctx.check(self.lib.vg_state_set(self.handle, view_slot, s.ctypes.data), "vg_state_set")

    def pool(self, buffer: int) -> np.ndarray:
        size = ctypes.c_uint64()
        self.ctx.check(self.lib.vg_scene_pool_read(self.handle, buffer, None, 0, ctypes.byref(size)), "vg_scene_pool_read")
        dtype = np.uint32 if buffer in (VG_POOL_MESHLET_INFO, VG_POOL_MESHLET_TRIANGLES, VG_POOL_MESHLET_VERTICES) else np.float32
        out = np.empty(size.value // 4, dtype=dtype)
        self.ctx.check(self.lib.vg_scene_pool_read(self.handle, buffer, out.ctypes.data, size.value, None), "vg_scene_pool_read")
        return out

    def profile(self, enable: bool = True):
        self.ctx.check(self.lib.vg_profile(self.handle, 1 if enable else 0), "vg_profile")

    def profile_read(self):
        """Per kernel position: (name, summed ms); and the number of timed launch sequences."""
        ms = (ctypes.c_double * 16)()
        names = (ctypes.c_char_p * 16)()
        seq = ctypes.c_uint64()
        self.ctx.check(self.lib.vg_profile_read(self.handle, ctypes.cast(ms, ctypes.c_void_p), 16, ctypes.byref(seq),
                                                ctypes.cast(names, ctypes.c_void_p)), "vg_profile_read")
        out = [(names[i].decode(), float(ms[i])) for i in range(16) if names[i]]
        return out, int(seq.value)

    def footprint(self, views):
        """(distinct instances tested, distinct meshlet records in candidate ranges) of one cut of `views`."""
        arr = (View * len(views))(*views)
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(self.lib.vg_footprint(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(views), ctypes.byref(a),
                                             ctypes.byref(b)), "vg_footprint")
        return int(a.value), int(b.value)

    @property
    def schedule(self) -> Schedule:
        sc = Schedule()
        self.ctx.check(self.lib.vg_schedule_get(self.handle, ctypes.byref(sc)), "vg_schedule_get")
        return sc

    def set_schedule(self, instance_blocks_per_cu: int, meshlet_blocks_per_cu: int, meshlet_waves_per_simd: int):
        sc = Schedule(instance_blocks_per_cu, meshlet_blocks_per_cu, meshlet_waves_per_simd)
        self.ctx.check(self.lib.vg_schedule_set(self.handle, ctypes.byref(sc)), "vg_schedule_set")

    def tune(self, views, repeats: int = 5):
        """Times candidate launch schedules on `views`, keeps the fastest; returns (schedule, median ms)."""
        arr = (View * len(views))(*views)
        sc, ms = Schedule(), ctypes.c_float()
        self.ctx.check(self.lib.vg_tune(self.handle, ctypes.cast(arr, ctypes.c_void_p), len(views), repeats, ctypes.byref(sc),
                                        ctypes.byref(ms)), "vg_tune")
        return sc, float(ms.value)

    @property
    def mesh_slots(self) -> int:
        return int(self.lib.vg_scene_mesh_slots(self.handle))

    def instance_base(self, slot: int):
        base, count = ctypes.c_uint32(), ctypes.c_uint32()
        self.ctx.check(self.lib.vg_mesh_instance_base(self.handle, slot, ctypes.byref(base), ctypes.byref(count)), "vg_mesh_instance_base")
        return int(base.value), int(count.value)

    @property
    def device_bytes(self) -> int:
        return int(self.lib.vg_scene_bytes(self.handle))

    def close(self):
        if self.handle:
            self.lib.vg_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request server over resident scenes (vg_server_*): continuous batching on `lanes` HIP streams."""

    def __init__(self, ctx: "Context", hbm_budget: int = 0, lanes: int = 4, max_batch_views: int = 0, max_batch_requests: int = 0,
                 max_bypass_us: int = 0, first_work_items: int = 0, policy: int = 0, max_blocks_per_cu: int = 0):
        self.ctx = ctx
        self.lib = ctx.lib
        self.scenes = []  # keeps the Scene objects alive while the server uses them
        cfg = ServerConfig(hbm_budget, lanes, max_batch_views, max_batch_requests, max_bypass_us, first_work_items, policy,
                           max_blocks_per_cu)
        handle = ctypes.c_void_p()
        ctx.check(self.lib.vg_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(handle)), "vg_server_create")
        self.handle = handle

    def check(self, rc: int, what: str):
        if rc != VG_OK:
            msg = self.lib.vg_server_last_error(self.handle)
            raise VgError(rc, f"{what}: {msg.decode() if msg else ''}")

    def add_scene(self, scene: "Scene") -> int:
        sid = ctypes.c_uint32()
        self.check(self.lib.vg_server_add_scene(self.handle, scene.handle, ctypes.byref(sid)), "vg_server_add_scene")
        self.scenes.append(scene)
        return int(sid.value)

    def submit(self, scene_id: int, views) -> int:
        arr = (View * len(views))(*views)
        t = ctypes.c_uint64()
        self.check(self.lib.vg_server_submit(self.handle, scene_id, ctypes.cast(arr, ctypes.c_void_p), len(views), ctypes.byref(t)),
                   "vg_server_submit")
        return int(t.value)

    def submit_array(self, scene_id: int, views_array, count: int) -> int:
        """submit() for a prebuilt ctypes array of View (no per-call copy of the records)."""
        t = ctypes.c_uint64()
        self.check(self.lib.vg_server_submit(self.handle, scene_id, ctypes.cast(views_array, ctypes.c_void_p), count, ctypes.byref(t)),
                   "vg_server_submit")
        return int(t.value)

    def poll(self, max_done: int = 256, wait_us: int = 0):
        done = (RequestDone * max_done)()
        n = ctypes.c_uint32()
        self.check(self.lib.vg_server_poll(self.handle, done, max_done, ctypes.byref(n), wait_us), "vg_server_poll")
        return [done[i] for i in range(n.value)]

    def read(self, ticket: int, view_index: int, max_entries: Optional[int] = None):
        """Counters, draw args, draw lists (camera, shadow) and metrics of one view of a completed request."""
        if max_entries is None:
            max_entries = self.ctx.config.max_meshlets
        counters = np.empty(self.ctx.counter_words, dtype=np.uint32)
        args = np.empty(self.ctx.draw_args_words, dtype=np.uint32)
        draws = np.empty((2, max_entries, 4), dtype=np.uint32)
        met = Metrics()
        self.check(self.lib.vg_server_read(self.handle, ticket, view_index, counters.ctypes.data, args.ctypes.data, draws.ctypes.data,
                                           max_entries, ctypes.byref(met)), "vg_server_read")
        return counters, args, draws, met

    def release(self, ticket: int):
        self.check(self.lib.vg_server_release(self.handle, ticket), "vg_server_release")

    def stats(self) -> ServerStats:
        st = ServerStats()
        self.check(self.lib.vg_server_stats_get(self.handle, ctypes.byref(st)), "vg_server_stats_get")
        return st

    def close(self):
        if self.handle:
            self.lib.vg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def budget_thresholds(lo: float, hi: float, bins: int) -> np.ndarray:
    """The budget histogram's threshold table (vg_budget_thresholds)."""
    out = np.empty(bins, dtype=np.float32)
    rc = load().vg_budget_thresholds(lo, hi, bins, out.ctypes.data)
    if rc != VG_OK:
        raise VgError(rc, "vg_budget_thresholds: need bins > 0 and 0 < lo <= hi")
    return out


def scan_probe(values: np.ndarray) -> np.ndarray:
    """Runs the device wave-scan probe on groups of 64 u32: per group (inclusive scan, group sum)."""
    lib = load()
    v = np.ascontiguousarray(values, dtype=np.uint32).reshape(-1, 64)
    out = np.empty((len(v), 2, 64), dtype=np.uint32)
    rc = lib.vg_scan_probe(v.ctypes.data, out.ctypes.data, len(v))
    if rc != VG_OK:
        raise VgError(rc, "vg_scan_probe failed")
    return out


def arith_probe(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Runs the device arithmetic probe: per pair (a / b, sqrt|a|, log|a|, rint(a), float(bits(b) >> 3))."""
    lib = load()
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.empty((len(a), 5), dtype=np.float32)
    rc = lib.vg_arith_probe(a.ctypes.data, b.ctypes.data, out.ctypes.data, len(a))
    if rc != VG_OK:
        raise VgError(rc, "vg_arith_probe failed")
    return out
