"""The expand pass on the GPU (k_expand through the C ABI) against the oracle's restatement of
expandPass (GeometryPool.ts:997-1021), index buffer for index buffer, bit-exact.

The expand pass is a function of a cut's own result (draw list, counters) and the pool: the oracle
expands the GPU's draw list, so the check holds also where the draw order is not determined (a draw
list that overflows). The cut itself is checked against the oracle's cut on the same views first.
"""
import numpy as np
import pytest

from helpers import compare_cut, ground_camera, load_gpu, small_ruins
from oracle import oracle
from vgcut import camera as cam
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu


def views_for(seed, count, shadow=True):
    rng = np.random.default_rng(seed)
    views = []
    for slot in range(count):
        if slot == count - 1:
            ortho = cam.orthographic_camera(-60, 60, 60, -60, 0.1, 500).look_at((30.0, 200.0, 20.0), (0.0, 0.0, 0.0))
            views.append(make_view(ortho, 2048, slot=slot))
            continue
        camera = ground_camera((rng.uniform(-50, 50), rng.uniform(2, 40), rng.uniform(-50, 50)),
                               (rng.uniform(-50, 50), 0.0, rng.uniform(-50, 50)))
        views.append(make_view(camera, [720, 1080, 2160][slot % 3], slot=slot, error_threshold=[1.0, 0.5, 2.0][slot % 3],
                               shadow_sweep=(20.0, -40.0, 0.0) if shadow else (0.0, 0.0, 0.0)))
    return views


def expand_and_check(loaded, views, label, graph=False, check_cut=True):
    scene = loaded.scene
    prev = {(v.slot, c): scene.read_indices(v.slot, c) for v in views for c in (0, 1)}
    refs = [oracle.cut(loaded.pool, v, state=scene.state(v.slot)) for v in views] if check_cut else None
    print(f"{label}: cut + expand of {len(views)} views", flush=True)
    scene.cut(views, graph=graph)
    scene.sync()
    totals = []
    for i, v in enumerate(views):
        gpu = scene.read(v.slot)
        counters, _, draws, _ = gpu
        if check_cut:
            compare_cut(gpu, refs[i], loaded.ctx.config.max_meshlets, f"{label} slot {v.slot}")
        for cut in (0, 1):
            got = scene.read_indices(v.slot, cut)
            want, written = oracle.expand(loaded.pool, cut, counters, draws, prev[(v.slot, cut)])
            diff = np.nonzero(got != want)[0]
            assert len(diff) == 0, (f"{label} slot {v.slot} cut {cut}: {len(diff)} indices differ, first at {diff[:8].tolist()}: "
                                    f"gpu {got[diff[:8]].tolist()} oracle {want[diff[:8]].tolist()}")
            requested = int(counters[2 + cut * 5 + 4])
            cap = int(loaded.pool.cut_capacity[cut])
            if requested <= cap:
                assert written == requested, f"{label} slot {v.slot} cut {cut}: {written} indices written of {requested}"
            else:
                assert written <= cap
            totals.append(written)
    return totals


def test_expand_batched_views_over_frames(gpu_available):
    loaded = load_gpu(small_ruins(seed=61, shadows=True), max_views=5, expand=True)
    views = views_for(61, 5)
    for frame in range(2):  # the second frame overwrites what the first wrote
        totals = expand_and_check(loaded, views, f"frame {frame}")
        assert sum(totals) > 0
    expand_and_check(loaded, views, "graph replay", graph=True)


def test_expand_clamped_by_index_capacity(gpu_available):
    """Index capacity below the demand: the controller retries coarser, then triangles past the capacity
    are skipped (index + 2 < capacity)."""
    loaded = load_gpu(small_ruins(seed=62, shadows=True), max_views=3, expand=True, max_triangles=30_000, max_shadow=10_000)
    views = views_for(62, 3)
    totals = expand_and_check(loaded, views, "clamped")
    assert sum(totals) > 0


def test_expand_with_draw_list_overflow(gpu_available):
    """A draw list that overflows (the draw order then decides which entries are written): the expand
    pass still equals the oracle's expansion of the GPU's own draw list."""
    loaded = load_gpu(small_ruins(seed=63, shadows=True), max_views=2, expand=True, max_meshlets=512)
    views = views_for(63, 2)
    totals = expand_and_check(loaded, views, "overflow", check_cut=False)
    assert sum(totals) > 0
