"""ORACLE — test infrastructure only: ctypes access to the C restatement (vgoracle.c).

Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this module, as the
checker. Inputs are the pool buffers in the reference layout (GeometryPool.ts:221-231) and a view
record (VirtualGeometry.ts:78-183).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "liboracle.so")


class OrScene(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("mesh_slots", ctypes.c_uint32),
        ("mesh_table", ctypes.c_void_p),
        ("instances", ctypes.c_void_p),
        ("instance_count", ctypes.c_uint32),
        ("cells", ctypes.c_void_p),
        ("cell_count", ctypes.c_uint32),
        ("bounds", ctypes.c_void_p),
        ("info", ctypes.c_void_p),
        ("meshlet_count", ctypes.c_uint32),
        ("max_meshlets", ctypes.c_uint32),
        ("cut_capacity", ctypes.c_uint32 * 2),
    ]


class OrView(ctypes.Structure):
    _fields_ = [
        ("view", ctypes.c_float * 16),
        ("proj", ctypes.c_float * 16),
        ("planes", ctypes.c_float * 24),
        ("sweep", ctypes.c_float * 3),
        ("proj_scale", ctypes.c_float),
        ("near_plane", ctypes.c_float),
        ("error_threshold", ctypes.c_float),
        ("shadow_error_scale", ctypes.c_float),
        ("lod_blend_max", ctypes.c_float),
        ("lod_scale_limit", ctypes.c_float),
        ("min_pixel_radius", ctypes.c_float),
        ("frustum_culling", ctypes.c_uint32),
        ("occlusion_active", ctypes.c_uint32),
        ("hzb", ctypes.c_void_p),
    ]


class OrStats(ctypes.Structure):
    _fields_ = [
        ("cells_tested", ctypes.c_uint64),
        ("instances_tested", ctypes.c_uint64),
        ("work_items", ctypes.c_uint64 * 2),
        ("clusters_tested", ctypes.c_uint64 * 2),
        ("occlusion_tests", ctypes.c_uint64),
        ("retries", ctypes.c_uint64 * 2),
        ("draw_overflow", ctypes.c_uint64 * 2),
        ("lod_checked", ctypes.c_uint64),
        ("near_threshold", ctypes.c_uint64),
        ("mismatch_near", ctypes.c_uint64),
        ("mismatch_hard", ctypes.c_uint64),
    ]


_lib = None


def build():
    """Compiles liboracle.so with the recipe in oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            build()
        _lib = _bind(ctypes.CDLL(LIBRARY))
    return _lib


def build_variant(out_path: str, mutant: int) -> str:
    """Compiles the oracle with one planted mistake (-DOR_MUTANT=mutant, see vgoracle.c) into out_path, for
    tests that show the oracle's checks catch such mistakes. Returns the path."""
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-D_GNU_SOURCE",
                    f"-DOR_MUTANT={int(mutant)}", "-shared", "-o", out_path, os.path.join(HERE, "vgoracle.c"), "-lm"], check=True)
    return out_path


def load_variant(path: str):
    """A ctypes handle on another build of the oracle (build_variant), bound like the default one."""
    return _bind(ctypes.CDLL(path))


def _bind(lib):
    if True:
        lib.or_counter_words.argtypes = [ctypes.c_uint32]
        lib.or_counter_words.restype = ctypes.c_uint32
        lib.or_cut.argtypes = [ctypes.POINTER(OrScene), ctypes.POINTER(OrView), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(OrStats)]
        lib.or_cut.restype = ctypes.c_int
        lib.or_hzb_size.argtypes = [ctypes.c_uint32]
        lib.or_hzb_size.restype = ctypes.c_uint32
        lib.or_hzb_build.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        lib.or_hzb_build.restype = ctypes.c_int
        lib.or_occluded.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]
        lib.or_occluded.restype = ctypes.c_int
        lib.or_logf.argtypes = [ctypes.c_float]
        lib.or_logf.restype = ctypes.c_float
        lib.or_budget_histogram.argtypes = [ctypes.POINTER(OrScene), ctypes.POINTER(OrView), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_uint32, ctypes.c_void_p]
        lib.or_budget_histogram.restype = ctypes.c_int
        lib.or_expand.argtypes = [ctypes.POINTER(OrScene), ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]
        lib.or_expand.restype = ctypes.c_int
        lib.or_cut_two_phase.argtypes = [ctypes.POINTER(OrScene), ctypes.POINTER(OrView), ctypes.POINTER(OrView), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_uint32, ctypes.POINTER(OrStats), ctypes.POINTER(OrStats), ctypes.c_void_p]
        lib.or_cut_two_phase.restype = ctypes.c_int
        lib.or_raster.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64,
                                  ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint32,
                                  ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        lib.or_raster.restype = ctypes.c_int
        P = ctypes.c_void_p
        lib.or_test_controller.argtypes = [ctypes.c_int, P, P, P, ctypes.c_uint32, ctypes.c_float, P, P]
        lib.or_test_controller.restype = ctypes.c_int
        lib.or_test_prefix.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, P, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint32, ctypes.c_uint32, P, P]
        lib.or_test_prefix.restype = ctypes.c_int
    return lib


@dataclass
class Pool:
    """Pool buffers in the reference layout (what the oracle reads)."""

    mesh_table: np.ndarray  # f32
    instances: np.ndarray  # f32, 20 per instance
    cells: np.ndarray  # f32, 8 per cell
    bounds: np.ndarray  # f32, 16 per meshlet
    info: np.ndarray  # u32, 4 per meshlet
    slots: int
    mesh_slots: int
    max_meshlets: int
    cut_capacity: tuple
    triangles: Optional[np.ndarray] = None  # u32 meshletTriangles of the pool
    vertices: Optional[np.ndarray] = None  # f32, 4 per vertex (x, y, z, 0)
    meshlet_vertices: Optional[np.ndarray] = None  # u32 pool-global vertex ids

    def struct(self) -> OrScene:
        s = OrScene()
        s.slots = self.slots
        s.mesh_slots = self.mesh_slots
        s.mesh_table = self.mesh_table.ctypes.data
        s.instances = self.instances.ctypes.data
        s.instance_count = len(self.instances) // 20
        s.cells = self.cells.ctypes.data
        s.cell_count = len(self.cells) // 8
        s.bounds = self.bounds.ctypes.data
        s.info = self.info.ctypes.data
        s.meshlet_count = len(self.info) // 4
        s.max_meshlets = self.max_meshlets
        s.cut_capacity[0] = self.cut_capacity[0]
        s.cut_capacity[1] = self.cut_capacity[1]
        return s


def pool_from_scene(scene, max_meshlets: int, max_triangles: int, max_shadow_triangles: int) -> Pool:
    """Reads a vgcut scene's host-side pool buffers (vg_scene_pool_read)."""
    from vgcut import native

    def f(b, dtype):
        a = scene.pool(b)
        return np.ascontiguousarray(a, dtype=dtype)

    return Pool(
        mesh_table=f(native.VG_POOL_MESH_TABLE, np.float32),
        instances=f(native.VG_POOL_INSTANCES, np.float32),
        cells=f(native.VG_POOL_CELLS, np.float32),
        bounds=f(native.VG_POOL_MESHLET_BOUNDS, np.float32),
        info=f(native.VG_POOL_MESHLET_INFO, np.uint32),
        slots=scene.ctx.config.mesh_slots,
        mesh_slots=scene.mesh_slots,
        max_meshlets=max_meshlets,
        cut_capacity=(max_triangles * 3, max_shadow_triangles * 3),
        triangles=f(native.VG_POOL_MESHLET_TRIANGLES, np.uint32),
        vertices=f(native.VG_POOL_VERTICES, np.float32),
        meshlet_vertices=f(native.VG_POOL_MESHLET_VERTICES, np.uint32),
    )


def view_struct(v, hzb: Optional[np.ndarray]) -> OrView:
    """An oracle view from a vgcut View (the same f32 uniforms)."""
    o = OrView()
    for i in range(16):
        o.view[i] = v.view_matrix[i]
        o.proj[i] = v.projection_matrix[i]
    for i in range(24):
        o.planes[i] = v.frustum_planes[i]
    for i in range(3):
        o.sweep[i] = v.shadow_sweep[i]
    o.proj_scale = v.proj_scale
    o.near_plane = v.near_plane
    o.error_threshold = v.error_threshold
    o.shadow_error_scale = v.shadow_error_scale
    o.lod_blend_max = v.lod_blend_max
    o.lod_scale_limit = v.lod_scale_limit
    o.min_pixel_radius = v.min_pixel_radius
    o.frustum_culling = v.frustum_culling
    o.occlusion_active = 1 if (v.occlusion and hzb is not None) else 0
    o.hzb = hzb.ctypes.data if hzb is not None else None
    return o


@dataclass
class CutResult:
    counters: np.ndarray
    args: np.ndarray
    draws: np.ndarray  # (2, max_meshlets, 4)
    state: np.ndarray
    stats: OrStats


def initial_state() -> np.ndarray:
    """Controller words at pool creation: scale 1 for both cuts (GeometryPool.ts:602-603)."""
    s = np.zeros(8, dtype=np.uint32)
    s[0] = s[1] = 10000
    return s


def cut(pool: Pool, view, state: Optional[np.ndarray] = None, hzb: Optional[np.ndarray] = None, fp64: bool = False,
        lib=None) -> CutResult:
    lib = lib or load()
    st = initial_state() if state is None else np.ascontiguousarray(state, dtype=np.uint32).copy()
    counters = np.zeros(lib.or_counter_words(pool.slots), dtype=np.uint32)
    args = np.zeros(pool.slots * 20, dtype=np.uint32)
    draws = np.zeros((2, pool.max_meshlets, 4), dtype=np.uint32)
    stats = OrStats()
    sc = pool.struct()
    hz = None if hzb is None else np.ascontiguousarray(hzb, dtype=np.float32)
    vs = view_struct(view, hz)
    rc = lib.or_cut(ctypes.byref(sc), ctypes.byref(vs), st.ctypes.data, counters.ctypes.data, args.ctypes.data, draws.ctypes.data,
                    1 if fp64 else 0, ctypes.byref(stats))
    if rc != 0:
        raise RuntimeError(f"or_cut failed ({rc})")
    return CutResult(np.concatenate([counters, st]), args, draws, st, stats)


@dataclass
class TwoPhaseResult:
    first: CutResult        # the cut (its draws array is `draws` below, with the second phase's entries after)
    counters2: np.ndarray   # second phase, reference layout (the 8 state words of the cut appended)
    args2: np.ndarray
    draws: np.ndarray       # (2, max_meshlets, 4)
    stats2: OrStats
    rejects: tuple          # (instances, meshlets, meshlet records dropped)


def cut_two_phase(pool: Pool, view, hzb1: Optional[np.ndarray], hzb2: Optional[np.ndarray], state: Optional[np.ndarray] = None,
                  reject_capacity: Optional[int] = None) -> TwoPhaseResult:
    """Two-phase occlusion: the cut against hzb1 recording its camera occlusion rejects, then the second
    phase against hzb2 (see or_cut_two_phase)."""
    lib = load()
    st = initial_state() if state is None else np.ascontiguousarray(state, dtype=np.uint32).copy()
    words = lib.or_counter_words(pool.slots)
    c1, c2 = np.zeros(words, dtype=np.uint32), np.zeros(words, dtype=np.uint32)
    a1, a2 = np.zeros(pool.slots * 20, dtype=np.uint32), np.zeros(pool.slots * 20, dtype=np.uint32)
    draws = np.zeros((2, pool.max_meshlets, 4), dtype=np.uint32)
    s1, s2 = OrStats(), OrStats()
    rejects = np.zeros(3, dtype=np.uint32)
    sc = pool.struct()
    h1 = None if hzb1 is None else np.ascontiguousarray(hzb1, dtype=np.float32)
    h2 = None if hzb2 is None else np.ascontiguousarray(hzb2, dtype=np.float32)
    v1, v2 = view_struct(view, h1), view_struct(view, h2)
    cap = pool.max_meshlets if reject_capacity is None else reject_capacity
    rc = lib.or_cut_two_phase(ctypes.byref(sc), ctypes.byref(v1), ctypes.byref(v2), st.ctypes.data, c1.ctypes.data, a1.ctypes.data,
                              c2.ctypes.data, a2.ctypes.data, draws.ctypes.data, cap, ctypes.byref(s1), ctypes.byref(s2),
                              rejects.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"or_cut_two_phase failed ({rc})")
    first = CutResult(np.concatenate([c1, st]), a1, draws, st, s1)
    return TwoPhaseResult(first, np.concatenate([c2, np.zeros(8, dtype=np.uint32)]), a2, draws, s2,
                          (int(rejects[0]), int(rejects[1]), int(rejects[2])))


def hzb_build(depth: np.ndarray, max_size: int = 1024) -> np.ndarray:
    lib = load()
    d = np.ascontiguousarray(depth, dtype=np.float32)
    h, w = d.shape
    out = np.zeros(lib.or_hzb_size(max_size), dtype=np.float32)
    rc = lib.or_hzb_build(d.ctypes.data, w, h, max_size, out.ctypes.data)
    if rc != 0:
        raise ValueError("or_hzb_build: depth larger than max_size")
    return out


def raster(pool: Pool, draws: np.ndarray, count: int, view_proj, width: int, height: int) -> np.ndarray:
    """The occluder raster of vg_hzb_from_cut: the first `count` entries of `draws` (camera entries, uvec4)
    drawn depth-only with `view_proj` (column-major) into a (height, width) image cleared to 1."""
    lib = load()
    d = np.ascontiguousarray(draws, dtype=np.uint32).reshape(-1, 4)
    assert count <= len(d)
    vp = np.ascontiguousarray(np.asarray(view_proj, dtype=np.float32).reshape(16))
    out = np.ones((height, width), dtype=np.float32)
    t = np.ascontiguousarray(pool.triangles if pool.triangles is not None else np.zeros(1, np.uint32), dtype=np.uint32)
    mv = np.ascontiguousarray(pool.meshlet_vertices if pool.meshlet_vertices is not None else np.zeros(1, np.uint32), dtype=np.uint32)
    vx = np.ascontiguousarray(pool.vertices if pool.vertices is not None else np.zeros(4, np.float32), dtype=np.float32)
    rc = lib.or_raster(pool.instances.ctypes.data, len(pool.instances) // 20, pool.info.ctypes.data, len(pool.info) // 4,
                       t.ctypes.data, len(t), mv.ctypes.data, len(mv), vx.ctypes.data, len(vx) // 4, d.ctypes.data, count,
                       vp.ctypes.data, width, height, out.ctypes.data)
    if rc != 0:
        raise ValueError("or_raster: an entry lies outside the pool")
    return out


def occluded(proj, near: float, hzb: np.ndarray, view_center, radius: float) -> bool:
    lib = load()
    p = np.ascontiguousarray(proj, dtype=np.float32)
    c = np.ascontiguousarray(view_center, dtype=np.float32)
    return bool(lib.or_occluded(p.ctypes.data, near, hzb.ctypes.data, c.ctypes.data, radius))


def logf(x: float) -> float:
    return float(load().or_logf(x))


def expand(pool: Pool, cut: int, counters: np.ndarray, draws: np.ndarray, indices: Optional[np.ndarray] = None, first: int = 0):
    """The expand pass (GeometryPool.ts:997-1021) over a cut result: `counters` of the view slot and its
    draw list `draws` (2 * max_meshlets uvec4, camera then shadow). Returns (indices, written): the index
    buffer of `cut` (a copy of `indices`, or zeros, with the written positions replaced) and the count."""
    lib = load()
    cap = int(pool.cut_capacity[cut])
    out = np.zeros(cap, dtype=np.uint32) if indices is None else np.array(indices, dtype=np.uint32, copy=True)
    assert len(out) == cap
    d = np.ascontiguousarray(draws, dtype=np.uint32).reshape(-1)
    assert len(d) == 2 * pool.max_meshlets * 4
    c = np.ascontiguousarray(counters, dtype=np.uint32)
    t = np.ascontiguousarray(pool.triangles, dtype=np.uint32)
    written = ctypes.c_uint64()
    sc = pool.struct()
    rc = lib.or_expand(ctypes.byref(sc), t.ctypes.data, len(t), cut, c.ctypes.data, d.ctypes.data, first, out.ctypes.data,
                       ctypes.byref(written))
    if rc != 0:
        raise RuntimeError(f"or_expand failed ({rc})")
    return out, int(written.value)


def budget_histogram(pool: Pool, view, thresholds: np.ndarray, state: Optional[np.ndarray] = None,
                     hzb: Optional[np.ndarray] = None) -> np.ndarray:
    lib = load()
    st = initial_state() if state is None else np.ascontiguousarray(state, dtype=np.uint32)
    t = np.ascontiguousarray(thresholds, dtype=np.float32)
    out = np.zeros(len(t), dtype=np.uint64)
    sc = pool.struct()
    hz = None if hzb is None else np.ascontiguousarray(hzb, dtype=np.float32)
    vs = view_struct(view, hz)
    rc = lib.or_budget_histogram(ctypes.byref(sc), ctypes.byref(vs), st.ctypes.data, t.ctypes.data, len(t), out.ctypes.data)
    if rc != 0:
        raise RuntimeError("or_budget_histogram failed")
    return out


def budget_pick(pool: Pool, view, budget: int, thresholds: np.ndarray, state: Optional[np.ndarray] = None,
                hzb: Optional[np.ndarray] = None):
    """The budget's choice for a view: (bin, per-bin triangles, the view at the chosen threshold).

    Bin b holds the triangles of the plain cut at thresholds[b]; the pick is the first bin at or under
    the budget, else the last. The chosen view's lod_scale_limit is max(1, maxErrorThreshold / t) in f32.
    """
    tris = budget_histogram(pool, view, thresholds, state, hzb)
    fits = np.nonzero(tris <= np.uint64(budget))[0]
    b = int(fits[0]) if len(fits) else len(thresholds) - 1
    chosen = type(view)()
    ctypes.memmove(ctypes.addressof(chosen), ctypes.addressof(view), ctypes.sizeof(view))
    t = np.float32(thresholds[b])
    chosen.error_threshold = float(t)
    chosen.lod_scale_limit = float(max(np.float32(1.0), np.float32(view.max_error_threshold) / t))
    chosen.triangle_budget = 0
    return b, tris, chosen


def controller(steer: bool, requested, selected, capacity, max_meshlets: int, lod_scale_limit: float,
               state: Optional[np.ndarray] = None, lib=None):
    """checkNode (steer False) or steerNode (steer True) of both cuts on given counts (GeometryPool.ts:1027-1069):
    (the 8 state words after it, [redo camera, redo shadow])."""
    lib = lib or load()
    st = initial_state() if state is None else np.ascontiguousarray(state, dtype=np.uint32).copy()
    req = np.ascontiguousarray(requested, dtype=np.uint32)
    sel = np.ascontiguousarray(selected, dtype=np.uint32)
    cap = np.ascontiguousarray(capacity, dtype=np.uint32)
    redo = np.zeros(2, dtype=np.int32)
    lib.or_test_controller(1 if steer else 0, req.ctypes.data, sel.ctypes.data, cap.ctypes.data, max_meshlets, lod_scale_limit,
                           st.ctypes.data, redo.ctypes.data)
    return st, redo.tolist()


def prefix(slots: int, counts, capacity: int, cut: int = 0, selected: int = 0, max_meshlets: int = 1 << 20, offset: int = 0,
           lib=None):
    """prefixPass of one cut (GeometryPool.ts:973-995) on per-(slot, part) index counts [[part 0, part 1], ...]:
    (counters, draw args of both cuts)."""
    lib = lib or load()
    c = np.ascontiguousarray(counts, dtype=np.uint32).reshape(-1)
    counters = np.zeros(lib.or_counter_words(slots), dtype=np.uint32)
    args = np.zeros(2 * slots * 2 * 5, dtype=np.uint32)
    rc = lib.or_test_prefix(slots, len(c) // 2, cut, c.ctypes.data, capacity, selected, max_meshlets, offset, counters.ctypes.data,
                            args.ctypes.data)
    if rc != 0:
        raise RuntimeError("or_test_prefix failed")
    return counters, args


def node_cpu_baseline(pool: Pool, views, scales=None, seconds: float = 12.0, threads: Optional[int] = None, timeout: float = 900.0):
    """Times the Node restatement of the reference's CPU cut (oracle/cpu_cut.js) on this pool and these views
    (vg_view records; `scales`: each view's camera LOD scale, 1 by default): its JSON result (CPU model, one
    thread over the views in order for about `seconds`, N worker threads over every view)."""
    import json
    import shutil
    import tempfile

    node = shutil.which("node")
    if node is None:
        raise RuntimeError("node is not installed")
    with tempfile.TemporaryDirectory(prefix="vgcut_cpu_") as d:
        for name, arr, dtype in (("mesh_table.f32", pool.mesh_table, np.float32), ("instances.f32", pool.instances, np.float32),
                                 ("cells.f32", pool.cells, np.float32), ("bounds.f32", pool.bounds, np.float32)):
            np.ascontiguousarray(arr, dtype=dtype).tofile(os.path.join(d, name))
        with open(os.path.join(d, "views.bin"), "wb") as f:
            for v in views:
                f.write(ctypes.string_at(ctypes.addressof(v), ctypes.sizeof(v)))
        s = np.ones(len(views), dtype=np.float32) if scales is None else np.asarray(scales, dtype=np.float32)
        s.tofile(os.path.join(d, "scales.f32"))
        cmd = [node, os.path.join(HERE, "cpu_cut.js"), d, "--seconds", str(seconds)]
        if threads:
            cmd += ["--threads", str(threads)]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"cpu_cut.js failed: {r.stderr[-2000:]}")
        return json.loads(r.stdout.strip().splitlines()[-1])
