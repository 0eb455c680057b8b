"""Triangle budget (our design for the reference's budget controller, VirtualGeometry.ts:549-588): the
projected-error histogram picks the smallest threshold whose plain cut fits the budget."""
import numpy as np
import pytest

from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenes
from vgcut.virtual_geometry import make_view


@pytest.fixture(scope="module")
def ruins_pool():
    spec = scenes.ruins(mesh_count=4, tri_range=(8_000, 40_000), instances=1500, half_extent=60.0, seed=23)
    ctx = native.Context(max_triangles=50_000_000, max_shadow_triangles=10_000_000, max_meshlets=1 << 16, device=-1)
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec)
    pool = oracle.pool_from_scene(scene, 1 << 16, 50_000_000, 10_000_000)
    yield pool
    scene.close()
    ctx.close()


def test_threshold_table():
    t = native.budget_thresholds(0.5, 3.0, 256)
    want = (0.5 * (3.0 / 0.5) ** (np.arange(256) / 255.0)).astype(np.float32)
    assert t[0] == np.float32(0.5) and t[-1] == np.float32(3.0)
    assert np.all(np.abs(t.view(np.int32) - want.view(np.int32)) <= 1) and np.all(np.diff(t) > 0)


def test_histogram_counts_the_plain_cut_exactly(ruins_pool):
    """With blending off, bin b's triangles equal the cut's requested triangles at thresholds[b]."""
    rng = np.random.default_rng(5)
    t = native.budget_thresholds(0.5, 3.0, 64)
    for _ in range(3):
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at(
            (rng.uniform(-40, 40), rng.uniform(2, 20), rng.uniform(-40, 40)), (rng.uniform(-50, 50), 0, rng.uniform(-50, 50)))
        v = make_view(camera, 1080, lod_blend=1.0, max_error_threshold=1e6, triangle_budget=1, threshold_range=(0.5, 3.0))
        tris = oracle.budget_histogram(ruins_pool, v, t)
        for b in (0, 17, 40, 63):
            vb = make_view(camera, 1080, error_threshold=float(t[b]), lod_blend=1.0, max_error_threshold=1e6)
            r = oracle.cut(ruins_pool, vb)
            assert int(r.counters[6]) == 3 * int(tris[b]), (b, int(r.counters[6]) // 3, int(tris[b]))
        assert tris[0] >= tris[-1] > 0


def test_pick_is_the_smallest_fitting_threshold(ruins_pool):
    camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at((0, 5, -50), (0, 0, 30))
    t = native.budget_thresholds(0.5, 3.0, 256)
    v = make_view(camera, 1080, triangle_budget=1, threshold_range=(0.5, 3.0))
    full = oracle.budget_histogram(ruins_pool, v, t)
    budget = int(full[128])
    v.triangle_budget = budget
    b, tris, chosen = oracle.budget_pick(ruins_pool, v, budget, t)
    assert tris[b] <= budget and (b == 0 or tris[b - 1] > budget)
    assert chosen.error_threshold == float(t[b])
