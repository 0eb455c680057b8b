"""GPU parity: the HIP cut through the C ABI against the oracle on the same inputs.

Bar: counters, controller state and draw args bit-exact; drawn-cluster sets (instance, meshlet,
triangles | blend range) identical after sorting; per (slot, part) the firstIndex ranges tile the
requested index count; work metrics (cells, instances, work items, cluster tests, occlusion tests)
equal.
"""
import math

import numpy as np
import pytest

from helpers import compare_cut, ground_camera, load_gpu, small_ruins, state_order_dependent, tiling_ok
from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenegen, scenes
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu


def test_device_arithmetic_matches_host(gpu_available):
    """Division, sqrt, the shared log, round-half-even and u32 -> f32 give the host's bits on gfx950."""
    rng = np.random.default_rng(1)
    n = 1 << 18
    a = (rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))).astype(np.float32)
    b = (rng.standard_normal(n) * np.exp(rng.uniform(-30, 30, n))).astype(np.float32)
    a[:8] = [0.5, 1.5, 2.5, -0.5, 1.0, 3.0, 1e-30, 2e30]
    b[b == 0] = 1.0
    out = native.arith_probe(a, b)
    with np.errstate(all="ignore"):
        assert np.array_equal(out[:, 0].view(np.uint32), (a / b).astype(np.float32).view(np.uint32)), "division"
        assert np.array_equal(out[:, 1].view(np.uint32), np.sqrt(np.abs(a)).view(np.uint32)), "sqrt"
    host_log = np.array([oracle.logf(float(x)) for x in np.abs(a[:4096])], dtype=np.float32)
    assert np.array_equal(out[:4096, 2].view(np.uint32), host_log.view(np.uint32)), "log"
    assert np.array_equal(out[:, 3], np.rint(a)), "rint"
    assert np.array_equal(out[:, 4], (b.view(np.uint32) >> 3).astype(np.float32)), "u32 -> f32"


def test_wave_scan_matches_cumsum(gpu_available):
    """The DPP wave scan behind every compaction and work-list offset: exact prefix sums and totals."""
    rng = np.random.default_rng(3)
    groups = 4096
    v = rng.integers(0, 1 << 20, (groups, 64), dtype=np.uint64).astype(np.uint32)
    v[0] = 0
    v[1] = 1
    v[2, ::7] = 0xFFFF
    v[3] = 0
    v[3, 63] = 5
    v[4] = 0
    v[4, 0] = 9
    mask = rng.random((groups, 64)) < 0.5
    v[8:] *= mask[8:]
    out = native.scan_probe(v)
    want = np.cumsum(v.astype(np.uint64), axis=1).astype(np.uint32)
    assert np.array_equal(out[:, 0], want)
    assert np.array_equal(out[:, 1], np.repeat(want[:, 63:64], 64, axis=1))


def _run(loaded, views, hzbs=None, fp64=False):
    loaded.scene.cut(views)
    loaded.scene.sync()
    results = []
    for v in views:
        gpu = loaded.scene.read(v.slot, max_entries=loaded.ctx.config.max_meshlets)
        results.append(gpu)
    return results


def test_icosphere_single_view(gpu_available):
    spec, camera, height = scenes.icosphere_view()
    loaded = load_gpu(spec, max_meshlets=1 << 20, max_triangles=10_000_000, max_shadow=4_000_000)
    view = make_view(camera, height)
    gpu, = _run(loaded, [view])
    ref = oracle.cut(loaded.pool, view, fp64=True)
    compare_cut(gpu, ref, 1 << 20, "icosphere")
    assert ref.stats.mismatch_hard == 0 and ref.stats.near_threshold == 0
    n = int(ref.counters[2])
    assert n > 300 and tiling_ok(gpu[2][0], n, gpu[0], 0, 64, loaded.slot_of_instance)  # 439 drawn


@pytest.mark.parametrize("threshold,shadows,seed", [(1.0, False, 3), (0.5, True, 4), (2.0, True, 5), (4.0, False, 6)])
def test_ruins_views(gpu_available, threshold, shadows, seed):
    """Many instances and meshes: frustum, draw distance, min pixel radius, shadow cut with a sweep, blend band."""
    spec = small_ruins(seed=seed, shadows=shadows)
    loaded = load_gpu(spec)
    rng = np.random.default_rng(seed)
    for k in range(3):
        eye = (rng.uniform(-40, 40), rng.uniform(1.5, 30), rng.uniform(-40, 40))
        target = (rng.uniform(-60, 60), 0.0, rng.uniform(-60, 60))
        camera = ground_camera(eye, target)
        sweep = (0.0, 0.0, 0.0) if not shadows else (30.0, -60.0, 10.0)
        view = make_view(camera, 1080, error_threshold=threshold, shadow_sweep=sweep)
        gpu, = _run(loaded, [view])
        state_before = oracle.initial_state() if k == 0 else state
        ref = oracle.cut(loaded.pool, view, state=state_before, fp64=True)
        state = ref.state
        compare_cut(gpu, ref, loaded.ctx.config.max_meshlets, f"ruins t={threshold} view {k}")
        assert ref.stats.mismatch_hard == 0
        for cut in (0, 1):
            n = min(int(ref.counters[2 + cut * 5]), loaded.ctx.config.max_meshlets)
            assert tiling_ok(gpu[2][cut], n, gpu[0], cut, 64, loaded.slot_of_instance)
        assert int(ref.counters[2]) > 0


def test_projective_instances(gpu_available):
    """Instance matrices whose last row is not (0, 0, 0, 1): the meshlet pass then keeps the transformed w of
    the LOD spheres (scenes of affine instances skip it, since it is exactly 1 there)."""
    spec = small_ruins(seed=8)
    rng = np.random.default_rng(8)
    m = spec.matrices[0].copy()
    m[::2, 3] = rng.uniform(-2e-3, 2e-3, len(m[::2])).astype(np.float32)  # c0.w: row 3 of column 0
    m[1::3, 15] = np.float32(1.25)  # c3.w
    spec.matrices[0] = m
    loaded = load_gpu(spec)
    cameras = (((-30.0, 6.0, -50.0), (10.0, 0.0, 40.0)), ((40.0, 20.0, 30.0), (-50.0, 0.0, -20.0)))
    # Every occupancy variant of the general meshlet pass (4 to 7 waves per SIMD).
    for sched in [(8, 8, 4), (6, 10, 5), (12, 12, 6), (12, 14, 7)]:
        loaded.scene.set_schedule(*sched)
        state = oracle.initial_state()
        loaded.scene.set_state(0, state)
        for k, (eye, target) in enumerate(cameras):
            view = make_view(ground_camera(eye, target), 1080, error_threshold=0.75)
            gpu, = _run(loaded, [view])
            ref = oracle.cut(loaded.pool, view, state=state, fp64=True)
            state = ref.state
            compare_cut(gpu, ref, loaded.ctx.config.max_meshlets, f"projective view {k}, schedule {sched}")
            assert ref.stats.mismatch_hard == 0 and int(ref.counters[2]) > 0


def test_capacity_retry_and_steer_over_frames(gpu_available):
    """Tiny draw buffers: overflowing cuts are redone coarser in the same frame; the scale glides across frames."""
    spec = small_ruins(seed=21, shadows=True)
    loaded = load_gpu(spec, max_meshlets=1 << 14, max_triangles=60_000, max_shadow=30_000)
    state = oracle.initial_state()
    retried = 0
    for f in range(6):
        camera = ground_camera((5.0 * f - 10, 3.0, -20.0), (0.0, 0.0, 40.0))
        view = make_view(camera, 1080, error_threshold=0.5, shadow_sweep=(10.0, -30.0, 0.0))
        gpu, = _run(loaded, [view])
        ref = oracle.cut(loaded.pool, view, state=state)
        state = ref.state
        retried += int(ref.stats.retries[0] + ref.stats.retries[1])
        if max(int(ref.counters[2]), int(ref.counters[7])) <= loaded.ctx.config.max_meshlets:
            compare_cut(gpu, ref, loaded.ctx.config.max_meshlets, f"capacity frame {f}")
        elif not state_order_dependent(ref):
            assert np.array_equal(gpu[0][-8:], ref.counters[-8:]), "controller state"
    assert retried > 0, "the test must exercise the retry rounds"


def test_occlusion_against_synthetic_hzb(gpu_available):
    """Two views, each with its own HZB from a rasterized noise heightfield; HZB pyramid equal to the oracle's."""
    spec = small_ruins(seed=31, instances=3000)
    loaded = load_gpu(spec, max_views=2, occlusion=True)
    views, hzbs = [], []
    for slot, (eye, target) in enumerate([((0.0, 2.0, -70.0), (0.0, 0.0, 60.0)), ((50.0, 2.0, 10.0), (-60.0, 2.0, 0.0))]):
        camera = ground_camera(eye, target)
        w, h = 480, 270
        vp = cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse)
        depth = scenegen.heightfield_depth(vp, w, h, (0.0, 0.0), 120.0, 128, 6.0, 1 / 25.0, 77 + slot)
        loaded.scene.set_hzb(slot, depth)
        hzb = oracle.hzb_build(depth, 1024)
        got = loaded.scene.hzb(slot)
        assert np.array_equal(got[:len(hzb)], hzb), "HZB pyramid differs from the oracle"
        views.append(make_view(camera, 540, occlusion=True, slot=slot))
        hzbs.append(hzb)
    results = _run(loaded, views)
    occluded = 0
    for v, gpu, hzb in zip(views, results, hzbs):
        ref = oracle.cut(loaded.pool, v, hzb=hzb)
        compare_cut(gpu, ref, loaded.ctx.config.max_meshlets, f"occlusion slot {v.slot}")
        occluded += int(ref.counters[1])
    assert occluded > 0, "the HZB must hide something"


def test_batched_views_equal_plain_launches(gpu_available):
    """One launch of 6 views == 6 single-view launches == the oracle per view (graph replay included)."""
    spec = small_ruins(seed=41, shadows=True)
    loaded = load_gpu(spec, max_views=6)
    rng = np.random.default_rng(41)
    views = []
    for slot in range(6):
        if slot < 4:
            camera = ground_camera((rng.uniform(-50, 50), rng.uniform(2, 40), rng.uniform(-50, 50)),
                                   (rng.uniform(-50, 50), 0.0, rng.uniform(-50, 50)))
            views.append(make_view(camera, [540, 720, 1080, 2160][slot], slot=slot, error_threshold=[1, 0.5, 2, 1][slot],
                                   shadow_sweep=(20.0, -40.0, 0.0)))
        else:
            ortho = cam.orthographic_camera(-60, 60, 60, -60, 0.1, 500).look_at((30.0, 200.0, 20.0), (0.0, 0.0, 0.0))
            views.append(make_view(ortho, 2048, slot=slot))
    plain = []
    for v in views:
        loaded.scene.cut([v])
        loaded.scene.sync()
        plain.append(loaded.scene.read(v.slot))
        # reset the slot's controller so the batched run starts from the same state
        loaded.scene.set_state(v.slot, oracle.initial_state())
    refs = [oracle.cut(loaded.pool, v) for v in views]
    for graph in (False, True, True):
        loaded.scene.cut(views, graph=graph)
        loaded.scene.sync()
        for v, p, ref in zip(views, plain, refs):
            got = loaded.scene.read(v.slot)
            compare_cut(got, ref, loaded.ctx.config.max_meshlets, f"batched slot {v.slot} graph={graph}")
            compare_cut(p, ref, loaded.ctx.config.max_meshlets, f"plain slot {v.slot}")
            loaded.scene.set_state(v.slot, oracle.initial_state())


def test_every_schedule_gives_the_same_cut(gpu_available):
    """Launch schedules (grids, meshlet-pass variants) change timing only: each equals the oracle, and the
    tuner restores the controller state and keeps a valid schedule."""
    spec = small_ruins(seed=45, shadows=True)
    loaded = load_gpu(spec, max_views=3, max_meshlets=1 << 20)  # index capacity overflows (retries), the draw list never
    views = [make_view(ground_camera((-40.0, 3.0, -40.0), (20.0, 0.0, 30.0)), 1080, slot=0, shadow_sweep=(10.0, -30.0, 5.0)),
             make_view(ground_camera((25.0, 30.0, 10.0), (-30.0, 0.0, -20.0)), 2160, slot=1, error_threshold=0.5),
             make_view(cam.orthographic_camera(-50, 50, 50, -50, 0.1, 400).look_at((10.0, 150.0, 0.0), (0.0, 0.0, 0.0)), 2048,
                       slot=2)]
    refs = [oracle.cut(loaded.pool, v) for v in views]
    assert sum(int(r.stats.retries[0] + r.stats.retries[1]) for r in refs) > 0
    assert not any(state_order_dependent(r) for r in refs)
    for sched in [(6, 10, 5), (1, 1, 4), (16, 32, 6), (3, 7, 5), (8, 4, 4), (12, 14, 7)]:
        loaded.scene.set_schedule(*sched)
        for graph in (False, True):
            loaded.scene.cut(views, graph=graph)
            loaded.scene.sync()
            for v, ref in zip(views, refs):
                compare_cut(loaded.scene.read(v.slot), ref, loaded.ctx.config.max_meshlets, f"schedule {sched} graph={graph} slot {v.slot}")
                loaded.scene.set_state(v.slot, oracle.initial_state())
    with pytest.raises(native.VgError):
        loaded.scene.set_schedule(6, 10, 8)
    before = [loaded.scene.state(v.slot).copy() for v in views]
    best, ms = loaded.scene.tune(views, repeats=2)
    assert ms > 0 and best.meshlet_waves_per_simd in (4, 5, 6, 7)
    assert loaded.scene.schedule.as_tuple() == best.as_tuple()
    for v, b in zip(views, before):
        assert np.array_equal(loaded.scene.state(v.slot), b)
    loaded.scene.cut(views)
    loaded.scene.sync()
    for v, ref in zip(views, refs):
        compare_cut(loaded.scene.read(v.slot), ref, loaded.ctx.config.max_meshlets, f"tuned {best.as_tuple()} slot {v.slot}")


def test_graph_replays_match_eager(gpu_available):
    """A captured launch sequence replayed several times, with eager work in between, equals eager runs."""
    spec = small_ruins(seed=51)
    loaded = load_gpu(spec)
    state = oracle.initial_state()
    for f in range(4):
        camera = ground_camera((-30.0 + 7 * f, 2.0, -30.0), (10.0, 0.0, 30.0))
        view = make_view(camera, 1080)
        loaded.scene.cut([view], graph=(f % 2 == 1))
        loaded.scene.sync()
        gpu = loaded.scene.read(0)
        ref = oracle.cut(loaded.pool, view, state=state)
        state = ref.state
        compare_cut(gpu, ref, loaded.ctx.config.max_meshlets, f"graph frame {f}")


@pytest.mark.parametrize("copies", [1, 6, 14])
def test_triangle_budget_pick_and_cut(gpu_available, copies):
    """Budget views pick the oracle's threshold on the GPU and cut exactly as the oracle does at it; a view
    without a budget in the same launch is unaffected. 3 and 18 views keep every view's histogram in the
    block's LDS; 42 views are past that limit, and each wave keeps its current view's histogram."""
    spec = small_ruins(seed=61, shadows=True)
    loaded = load_gpu(spec, max_views=3 * copies, max_triangles=20_000_000, max_shadow=10_000_000)
    bins = loaded.ctx.config.budget_bins
    t = native.budget_thresholds(0.5, 3.0, bins)
    views = []
    poses = [((0.0, 4.0, -60.0), (0.0, 0.0, 40.0), 300_000), ((30.0, 15.0, 20.0), (-40.0, 0.0, -10.0), 50_000),
             ((-20.0, 3.0, 0.0), (40.0, 0.0, 5.0), 0)]
    for slot in range(3 * copies):
        eye, target, budget = poses[slot % 3]
        k = slot // 3
        eye = (eye[0] + 3.0 * k, eye[1], eye[2] - 2.0 * k)
        camera = ground_camera(eye, target)
        views.append(make_view(camera, 1080, slot=slot, triangle_budget=budget * (1 + k) // 1, threshold_range=(0.5, 3.0),
                               shadow_sweep=(10.0, -30.0, 0.0)))
    for graph in (False, True):
        for v in views:
            loaded.scene.set_state(v.slot, oracle.initial_state())
        loaded.scene.cut(views, graph=graph)
        loaded.scene.sync()
        for v in views:
            gpu = loaded.scene.read(v.slot)
            if v.triangle_budget:
                b, tris, chosen = oracle.budget_pick(loaded.pool, v, v.triangle_budget, t)
                assert gpu[3].budget_bin == b and gpu[3].error_threshold == np.float32(t[b]), (gpu[3].budget_bin, b)
                ref = oracle.cut(loaded.pool, chosen)
            else:
                ref = oracle.cut(loaded.pool, v)
            compare_cut(gpu, ref, loaded.ctx.config.max_meshlets, f"budget slot {v.slot} graph={graph}")


def test_graph_replay_after_instance_changes(gpu_available):
    """A captured launch sequence is not replayed after the scene's contents change without a reallocation:
    a mesh set to no instances, then another mesh's instance count changed on a scene small enough that its
    cell buffers stay within their first allocation. Each graph cut equals the eager cut and the oracle."""
    from vgcut import scenes

    spec = small_ruins(seed=57, instances=300, meshes=3)
    loaded = load_gpu(spec)
    scene = loaded.scene
    view = make_view(ground_camera((-25.0, 6.0, -30.0), (10.0, 0.0, 25.0)), 1080)
    rng = np.random.default_rng(57)

    def check(label):
        scene.set_state(0, oracle.initial_state())
        scene.cut([view], graph=True)
        scene.sync()
        graph = scene.read(0)
        scene.set_state(0, oracle.initial_state())
        scene.cut([view])
        scene.sync()
        eager = scene.read(0)
        pool = oracle.pool_from_scene(scene, loaded.ctx.config.max_meshlets, 2_000_000, 1_000_000)
        ref = oracle.cut(pool, view)
        compare_cut(graph, ref, loaded.ctx.config.max_meshlets, f"{label}: graph")
        compare_cut(eager, ref, loaded.ctx.config.max_meshlets, f"{label}: eager")
        return int(ref.counters[2])

    drawn = [check("as loaded")]
    scene.set_instances(0, np.zeros((0, 16), dtype=np.float32))
    drawn.append(check("mesh 0 without instances"))
    n = len(spec.matrices[1])
    m = n + 140  # one more cell
    more = scenes.yaw_scale_translate(rng.uniform(0, 6.28, m), rng.uniform(0.5, 2.0, m),
                                      np.stack([rng.uniform(-30, 30, m), np.zeros(m), rng.uniform(-30, 30, m)], 1))
    scene.set_instances(1, more)
    drawn.append(check("mesh 1 with 140 more instances"))
    assert drawn[0] != drawn[1] and drawn[1] != drawn[2], drawn


def test_cut_after_mesh_removals(gpu_available):
    """VirtualGeometry.unregister (GeometryPool.remove): after a mesh is removed, its slot refilled by another
    mesh and instance ranges repacked, graph and eager cuts equal the oracle's cut of the pool as it stands,
    and the removed mesh's instances are no longer drawn."""
    from vgcut import scenes

    spec = small_ruins(seed=59, instances=600, meshes=3)
    loaded = load_gpu(spec)
    scene = loaded.scene
    view = make_view(ground_camera((-25.0, 6.0, -30.0), (10.0, 0.0, 25.0)), 1080)
    rng = np.random.default_rng(59)

    def check(label):
        scene.set_state(0, oracle.initial_state())
        scene.cut([view], graph=True)
        scene.sync()
        graph = scene.read(0)
        scene.set_state(0, oracle.initial_state())
        scene.cut([view])
        scene.sync()
        eager = scene.read(0)
        pool = oracle.pool_from_scene(scene, loaded.ctx.config.max_meshlets, 2_000_000, 1_000_000)
        ref = oracle.cut(pool, view)
        compare_cut(graph, ref, loaded.ctx.config.max_meshlets, f"{label}: graph")
        compare_cut(eager, ref, loaded.ctx.config.max_meshlets, f"{label}: eager")
        return graph

    check("as loaded")
    scene.remove_mesh(1)
    after = check("mesh 1 removed")
    n = int(after[0][2])
    slots = scene.pool(native.VG_POOL_INSTANCES).reshape(-1, 20)[:, 19]
    assert not np.any(slots[after[2][0][:n, 0]] == 1), "a removed mesh's instance was drawn"
    slot = scene.add_mesh(spec.meshes[2])
    assert slot == 1
    m = 500
    mats = scenes.yaw_scale_translate(rng.uniform(0, 6.28, m), rng.uniform(0.5, 2.0, m),
                                      np.stack([rng.uniform(-30, 30, m), np.zeros(m), rng.uniform(-30, 30, m)], 1))
    scene.set_instances(slot, mats)
    check("slot 1 refilled")
    scene.remove_mesh(0)
    scene.remove_mesh(2)
    check("only the refilled slot left")
