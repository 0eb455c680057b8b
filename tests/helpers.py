"""Shared test helpers: small scenes, GPU runs and the comparison with the oracle."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenegen, scenes
from vgcut.virtual_geometry import make_view


def small_ruins(seed=11, meshes=5, instances=2500, half_extent=80.0, tri_range=(8_000, 60_000), shadows=False):
    spec = scenes.ruins(mesh_count=meshes, tri_range=tri_range, instances=instances, half_extent=half_extent, seed=seed)
    if shadows:
        spec.cast_shadow = [k % 2 == 0 for k in range(meshes)]
    rng = np.random.default_rng(seed + 50)
    spec.max_draw_distance = [float(rng.uniform(40, 150)) if k % 3 == 1 else 1e30 for k in range(meshes)]
    spec.min_pixel_radius = [float(rng.uniform(0, 3)) if k % 2 == 0 else 0.0 for k in range(meshes)]
    return spec


def ground_camera(eye, target, fov=60.0, aspect=16 / 9, near=0.1, far=2000.0):
    return cam.perspective_camera(fov, aspect, near, far).look_at(eye, target)


def draw_set(draws_cut: np.ndarray, n: int):
    """Sorted (instance, meshlet, packed triangles | blend range) of the first n entries."""
    d = draws_cut[:n]
    keys = d[:, [0, 1, 3]]
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    return keys[order]


def compare_cut(gpu, ref: oracle.CutResult, max_meshlets: int, label: str = ""):
    """Bit-exact comparison with the oracle. The controller words are skipped when a round of the cut
    overflowed the draw list: its index counts then depend on the append order (in the reference too)."""
    counters, args, draws, metrics = gpu
    words = len(counters) - 8 if state_order_dependent(ref) else len(counters)
    diff = np.nonzero(counters[:words] != ref.counters[:words])[0]
    assert len(diff) == 0, (f"{label}: counters differ at {diff[:16].tolist()}: gpu {counters[diff[:16]].tolist()} "
                            f"oracle {ref.counters[diff[:16]].tolist()}")
    assert np.array_equal(args, ref.args), f"{label}: draw args differ"
    for cut in (0, 1):
        selected = int(ref.counters[2 + cut * 5])
        n = min(selected, max_meshlets)
        g = draw_set(draws[cut], n)
        r = draw_set(ref.draws[cut], n)
        assert np.array_equal(g, r), f"{label}: cut {cut} drawn sets differ ({n} entries)"
    st = ref.stats
    assert metrics.cells_tested == st.cells_tested, f"{label}: cells tested"
    assert metrics.instances_tested == st.instances_tested, f"{label}: instances tested"
    assert list(metrics.work_items) == list(st.work_items), f"{label}: work items"
    assert list(metrics.clusters_tested) == list(st.clusters_tested), f"{label}: clusters tested"
    assert metrics.occlusion_tests == st.occlusion_tests, f"{label}: occlusion tests"


def state_order_dependent(ref: oracle.CutResult) -> bool:
    return int(ref.stats.draw_overflow[0]) + int(ref.stats.draw_overflow[1]) > 0


def tiling_ok(draws_cut: np.ndarray, n: int, counters: np.ndarray, cut: int, slots: int, slot_of_instance) -> bool:
    """Per (mesh slot, part): firstIndex ranges of the drawn meshlets are disjoint and cover [0, requested)."""
    d = draws_cut[:n].astype(np.int64)
    if n == 0:
        return True
    tris = d[:, 3] & 255
    part = np.where((d[:, 3] & 0xFFFF00) == (255 << 16), 0, 1)
    mslot = slot_of_instance[d[:, 0]]
    for s in np.unique(mslot):
        for p in (0, 1):
            sel = (mslot == s) & (part == p)
            req = int(counters[12 + cut * slots * 4 + s * 4 + p * 2])
            if not sel.any():
                if req != 0:
                    return False
                continue
            start = d[sel, 2]
            length = tris[sel] * 3
            order = np.argsort(start)
            start, length = start[order], length[order]
            if start[0] != 0 or np.any(start[1:] != start[:-1] + length[:-1]) or start[-1] + length[-1] != req:
                return False
    return True


@dataclass
class Loaded:
    scene: native.Scene
    ctx: native.Context
    pool: oracle.Pool
    slot_of_instance: np.ndarray


def load_gpu(spec, max_meshlets=1 << 16, max_triangles=2_000_000, max_shadow=1_000_000, max_views=1, occlusion=False, mesh_slots=64,
             expand=False):
    ctx = native.Context(max_triangles=max_triangles, max_shadow_triangles=max_shadow, max_meshlets=max_meshlets,
                         max_views=max_views, mesh_slots=mesh_slots, device=0, expand=expand)
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=occlusion)
    pool = oracle.pool_from_scene(scene, max_meshlets, max_triangles, max_shadow)
    slot_of_instance = pool.instances.reshape(-1, 20)[:, 19].astype(np.int64)
    return Loaded(scene, ctx, pool, slot_of_instance)
