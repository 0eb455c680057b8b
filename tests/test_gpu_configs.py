"""Parity at the benchmark configurations' full sizes (through the C ABI, against the oracle):

- the ruins scene (113,106 instances of 48 props of 0.2M-5M triangles): 8 seeded poses at 2 m eye
  height, each at 1080p and 2160p, occlusion against the rasterized 30 m heightfield, the reference's
  default triangle capacities (so the capacity retries run) with a draw list of 8M entries, all 16 views
  in one launch. No round may overflow the draw list: an overflowed list keeps the entries appended
  first, so the requested-triangle counts and the retry's scale depend on the append order (in the
  reference too);
- the 3M-instance world with its terrain as a mesh (config 2): every fourth frame of the 16-frame seeded fly-through
  (4 frames; the oracle's budget histogram takes about 20 s per frame; bench.py --plain-configs times all
  16) at 1080p with an 8M-triangle budget (histogram pick; this world's cut exceeds 8M triangles even at
  3 px, so the pick is clamped to the last threshold of thresholdRange, as the reference clamps) and
  two-phase occlusion against the previous frame's visible set: the first phase's HZB is the occluder
  raster of the previous frame's cut (both phases) with this frame's camera, the second phase's the raster
  of the first phase's clusters, each HZB equal to the oracle's raster of the oracle's own cut; the
  controller state carried over; the second phases must recover clusters. Then frame 0's camera again
  with a budget of the plain cut's triangles at the middle threshold, so the histogram's pick lands inside
  the table at full size;

- the same world cut for its 64 mixed views (config 3: main cameras at 1080p and 2160p with terrain
  occlusion, orthographic shadow cascades, cube-map faces, 720p viewers) in one graph-replayed launch,
  and the same views in 64 plain one-view launches from the same controller state: batched, plain and
  the oracle agree view by view;
- served requests over the 8 resident scenes of 12k-3M instances (config 4): a seeded mix submitted at
  once to a server with 4 lanes; each request's result equals its result alone (one request per launch),
  and for one request per scene, the oracle's cut of each of its views.

Counters (controller state included), draw args and drawn sets are compared bit for bit, as everywhere.
"""
import math
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import compare_cut, state_order_dependent
from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenegen, scenes
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu


def sorted_keys(d):
    k = d[:, [0, 1, 3]]
    return k[np.lexsort(k.T[::-1])]


def terrain_depth(spec, camera, w, h, extent):
    hf = spec.heightfield
    vp = cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse)
    return scenegen.heightfield_depth(vp, w, h, (0.0, 0.0), extent, 256, hf["amplitude"], hf["frequency"], hf["seed"])


def test_ruins_full_size_16_views(gpu_available):
    spec = scenes.ruins()
    assert spec.instance_count == 113_106 and len(spec.meshes) == 48
    ctx = native.Context(max_meshlets=8 << 20, max_views=16, device=0)  # 10M / 4M triangles: the reference's defaults
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=True)
    pool = oracle.pool_from_scene(scene, ctx.config.max_meshlets, ctx.config.max_triangles, ctx.config.max_shadow_triangles)
    rng = np.random.default_rng(4)
    views, hzbs = [], []
    for pose in range(8):
        eye = (float(rng.uniform(-400, 400)), 2.0, float(rng.uniform(-400, 400)))
        yaw = float(rng.uniform(0, 2 * math.pi))
        target = (eye[0] + 10 * math.cos(yaw), 2.0, eye[2] + 10 * math.sin(yaw))
        for k, (height, hw, hh) in enumerate(((1080, 960, 540), (2160, 1024, 576))):
            camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at(eye, target)
            slot = pose * 2 + k
            depth = terrain_depth(spec, camera, hw, hh, 500.0)
            scene.set_hzb(slot, depth)
            views.append(make_view(camera, height, slot=slot, occlusion=True))
            hzbs.append(oracle.hzb_build(depth, 1024))
    print("ruins: one launch of 16 views", flush=True)
    scene.cut(views)
    scene.sync()
    occluded = 0
    for i, (v, hzb) in enumerate(zip(views, hzbs)):
        ref = oracle.cut(pool, v, hzb=hzb, fp64=i < 2)
        assert ref.stats.mismatch_hard == 0
        assert not state_order_dependent(ref), f"ruins view {v.slot}: the draw list overflowed"
        compare_cut(scene.read(v.slot), ref, ctx.config.max_meshlets, f"ruins view {v.slot}")
        occluded += int(ref.counters[1])
    assert occluded > 0


def vp_of(camera):
    return np.asarray(cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse), dtype=np.float32)


@pytest.fixture(scope="module")
def world_spec():
    """The 3M-instance world with its terrain (64 props DAG-built at the reference's defaults), built once."""
    print("building the world scene", flush=True)
    return scenes.world(terrain_grid=512)


def test_world_flythrough_budget_two_phase(gpu_available, world_spec):
    spec = world_spec
    assert spec.instance_count == 3_000_001  # the props and the terrain
    ctx = native.Context(max_triangles=20_000_000, max_shadow_triangles=10_000_000, max_meshlets=4 << 20, max_views=1,
                         mesh_slots=72, device=0)
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=True)
    pool = oracle.pool_from_scene(scene, ctx.config.max_meshlets, ctx.config.max_triangles, ctx.config.max_shadow_triangles)
    mm = ctx.config.max_meshlets
    budget = 8_000_000
    thresholds = native.budget_thresholds(0.5, 3.0, ctx.config.budget_bins)
    path = scenes.flythrough(spec, 16, 6144.0, 30.0, seed=9)
    state = oracle.initial_state()
    prev_draws, prev_n = None, 0
    added = hidden = 0
    w, h = 960, 540
    first = None  # frame 0's view and histogram, for the interior pick below
    for f in range(0, 16, 4):
        eye, target = path[f]
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 20000).look_at(eye, target)
        vp = vp_of(camera)
        # The first phase's occluders: the previous frame's visible set (both phases) drawn with this camera.
        if prev_draws is None:
            far = np.ones((h, w), dtype=np.float32)
            scene.set_hzb(0, far)
            hzb1 = oracle.hzb_build(far, 1024)
        else:
            scene.hzb_from_cut(0, vp, w, h, with_phase2=True)
            hzb1 = oracle.hzb_build(oracle.raster(pool, prev_draws, prev_n, vp, w, h), 1024)
            assert np.array_equal(scene.hzb(0)[:len(hzb1)], hzb1), f"frame {f}: first-phase HZB differs"
        view = make_view(camera, 1080, occlusion=2, triangle_budget=budget, threshold_range=(0.5, 3.0))
        b, tris, chosen = oracle.budget_pick(pool, view, budget, thresholds, state=state, hzb=hzb1)
        if first is None:
            first = (camera, tris)
        ref1 = oracle.cut(pool, chosen, state=state, hzb=hzb1)
        print(f"world frame {f}: budget bin {b}", flush=True)
        scene.cut([view])
        scene.sync()
        gpu = scene.read(0)
        assert gpu[3].budget_bin == b, (f, gpu[3].budget_bin, b)
        assert not state_order_dependent(ref1), f"world frame {f}: the draw list overflowed"
        compare_cut(gpu, ref1, mm, f"world frame {f} first phase")
        # The second phase's occluders: this frame's first-phase clusters.
        n1 = min(int(ref1.counters[2]), mm)
        scene.hzb_from_cut(0, vp, w, h)
        hzb2 = oracle.hzb_build(oracle.raster(pool, ref1.draws[0], n1, vp, w, h), 1024)
        assert np.array_equal(scene.hzb(0)[:len(hzb2)], hzb2), f"frame {f}: second-phase HZB differs"
        ref = oracle.cut_two_phase(pool, chosen, hzb1, hzb2, state=state)
        scene.cut_phase2([view])
        scene.sync()
        c2, a2 = scene.read_phase2(0)
        assert np.array_equal(c2, ref.counters2), f"frame {f}: second-phase counters differ"
        assert np.array_equal(a2, ref.args2), f"frame {f}: second-phase args differ"
        draws = scene.read(0)[2]
        n2 = min(int(c2[2]), mm)
        assert np.array_equal(sorted_keys(draws[0, n1:n2]), sorted_keys(ref.draws[0, n1:n2])), f"frame {f}: second-phase sets differ"
        added += n2 - n1
        hidden += int(ref1.counters[1])
        state = ref.first.state
        prev_draws, prev_n = ref.draws[0], n2
    print(f"world fly-through: {hidden} instances hidden by the first phases, {added} clusters added by the second", flush=True)
    assert hidden > 0, "the previous frames' visible sets must hide something"
    assert added > 0, "the second phases must recover clusters the moving camera uncovers"

    # An interior pick at full size: frame 0's camera from the controller at rest with a budget of the plain
    # cut's triangles at the middle threshold, so the pick lands strictly inside the table.
    camera, tris = first
    bins = len(thresholds)
    mid = int(tris[bins // 2])
    expect = int(np.nonzero(tris <= np.uint64(mid))[0][0])
    assert 0 < expect < bins - 1 and tris[0] > mid, (expect, tris[0], mid)
    far = np.ones((h, w), dtype=np.float32)
    scene.set_hzb(0, far)
    scene.set_state(0, oracle.initial_state())
    view = make_view(camera, 1080, occlusion=1, triangle_budget=mid, threshold_range=(0.5, 3.0))
    b, _, chosen = oracle.budget_pick(pool, view, mid, thresholds, state=oracle.initial_state(), hzb=oracle.hzb_build(far, 1024))
    assert b == expect
    ref = oracle.cut(pool, chosen, state=oracle.initial_state(), hzb=oracle.hzb_build(far, 1024))
    scene.cut([view])
    scene.sync()
    gpu = scene.read(0)
    assert gpu[3].budget_bin == b, (gpu[3].budget_bin, b)
    assert not state_order_dependent(ref), "the interior pick's draw list overflowed"
    compare_cut(gpu, ref, mm, "world frame 0 interior budget pick")


def compact(counters, args, draws, metrics, mm):
    """(counters, args, sorted drawn sets of both cuts, work metrics) of one view's result."""
    sets = [sorted_keys(draws[c, :min(int(counters[2 + 5 * c]), mm)]) for c in (0, 1)]
    met = (int(metrics.cells_tested), int(metrics.instances_tested), tuple(metrics.work_items), tuple(metrics.clusters_tested),
           int(metrics.occlusion_tests)) if metrics is not None else None
    return counters.copy(), args.copy(), sets, met


def read_compact(read, mm):
    """A view's result read with a draw list just long enough for what it holds: `read(max_entries)` is
    Scene.read or Server.read bound to the view."""
    counters = read(1)[0]
    n = max(1, min(max(int(counters[2]), int(counters[7])), mm))
    c, a, d, m = read(n)
    return compact(c, a, d, m, mm)


def oracle_compact(ref, mm):
    st = ref.stats
    met = (int(st.cells_tested), int(st.instances_tested), tuple(st.work_items), tuple(st.clusters_tested), int(st.occlusion_tests))
    return compact(ref.counters, ref.args, ref.draws, None, mm)[:3] + (met,)


def differ(a, b):
    """Where two compact results differ (empty when they are the same)."""
    out = []
    d = np.nonzero(a[0] != b[0])[0]
    if len(d):
        out.append(f"counters at {d[:8].tolist()}: {a[0][d[:8]].tolist()} vs {b[0][d[:8]].tolist()}")
    if not np.array_equal(a[1], b[1]):
        out.append("draw args")
    for c in (0, 1):
        if not np.array_equal(a[2][c], b[2][c]):
            out.append(f"cut {c} drawn sets ({len(a[2][c])} vs {len(b[2][c])} entries)")
    if a[3] is not None and b[3] is not None and a[3] != b[3]:
        out.append(f"work metrics {a[3]} vs {b[3]}")
    return "; ".join(out)


def test_world_64_mixed_views_batched_plain_oracle(gpu_available, world_spec):
    spec = world_spec
    mm = 1 << 23
    ctx = native.Context(max_triangles=300_000_000, max_shadow_triangles=100_000_000, max_meshlets=mm, max_views=64, mesh_slots=72,
                         device=0)
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=True)
    vd = scenes.mixed_views(spec, 64)
    views = [v for v, _ in vd]
    for v, depth in vd:
        if depth is not None:
            scene.set_hzb(v.slot, depth)
    print("world: one graph launch of 64 mixed views", flush=True)
    scene.cut(views, graph=True)
    scene.sync()
    batched = [read_compact(lambda n, s=v.slot: scene.read(s, n), mm) for v in views]
    for v in views:
        scene.set_state(v.slot, oracle.initial_state())
    print("world: 64 plain launches", flush=True)
    for v in views:
        scene.cut([v])
    scene.sync()
    plain = [read_compact(lambda n, s=v.slot: scene.read(s, n), mm) for v in views]
    pool = oracle.pool_from_scene(scene, mm, ctx.config.max_triangles, ctx.config.max_shadow_triangles)
    scene.close()
    ctx.close()
    hzbs = {v.slot: oracle.hzb_build(d, 1024) for v, d in vd if d is not None}

    def reference(v):
        ref = oracle.cut(pool, v, hzb=hzbs.get(v.slot))
        return oracle_compact(ref, mm), state_order_dependent(ref), int(ref.stats.mismatch_hard)

    print("world: the oracle's cut of every view", flush=True)
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=16) as ex:
        refs = list(ex.map(reference, views))
    print(f"world: oracle done in {time.time() - t0:.1f} s", flush=True)
    tested = 0
    for v, b, p, (r, over, hard) in zip(views, batched, plain, refs):
        assert not over, f"view {v.slot}: the draw list overflowed"
        assert hard == 0
        assert not differ(b, r), f"view {v.slot} batched vs oracle: {differ(b, r)}"
        assert not differ(p, r), f"view {v.slot} plain vs oracle: {differ(p, r)}"
        tested += sum(r[3][3])
    assert tested > 5_000_000, tested  # the full workload: every view's cluster tests (8.8M on this world)


def test_serving_8_resident_scenes(gpu_available, world_spec):
    props = world_spec.meshes[:world_spec.props]
    # Draw lists of the largest size (16M entries per cut): a round whose draw list overflows appends in an
    # order the launch decides, so the requests compared must not overflow in any round.
    mm = 1 << 24
    ctx = native.Context(max_triangles=100_000_000, max_shadow_triangles=10_000_000, max_meshlets=mm, max_views=16, mesh_slots=64,
                         device=0)
    print("serving: building 8 resident scenes", flush=True)
    specs = scenes.serving_specs(props)
    resident = []
    for sp, _ in specs:
        scene = native.Scene(ctx)
        scenes.load_into(scene, sp)
        resident.append(scene)
    assert [sp.instance_count for sp, _ in specs] == list(scenes.SERVE_SIZES)
    requests = scenes.serve_requests(specs, 40, seed=12)
    assert len({k for k, _, _ in requests}) == 8, "the mix must hold every scene"

    def run(server, ids, order):
        """Submits the requests `order` at once; reads each result as it completes and releases it (which
        frees view slots for requests still queued). Returns {request: per-view results}, {request: batch size}."""
        out, batches, tickets = {}, {}, {}
        for i in order:
            k, arr, n = requests[i]
            tickets[server.submit_array(ids[k], arr, n)] = i
        t0 = time.time()
        pending = set(tickets)
        while pending:
            for d in server.poll(64, wait_us=2000):
                i = tickets[d.ticket]
                assert d.status == native.VG_OK, (i, d.status)
                out[i] = [read_compact(lambda m, j=j, t=d.ticket: server.read(t, j, m), mm) for j in range(requests[i][2])]
                batches[i] = d.batch_requests
                server.release(d.ticket)
                pending.discard(d.ticket)
            assert time.time() - t0 < 120, f"{len(pending)} requests did not complete"
        return out, batches

    print("serving: each request alone", flush=True)
    alone_server = native.Server(ctx, lanes=1, max_batch_views=16, max_batch_requests=1)
    alone_ids = [alone_server.add_scene(sc) for sc in resident]
    alone = {}
    for i in range(len(requests)):
        got, _ = run(alone_server, alone_ids, [i])
        alone.update(got)
    alone_server.close()
    print("serving: the mix on 4 lanes", flush=True)
    server = native.Server(ctx, lanes=4, max_batch_views=16)
    ids = [server.add_scene(sc) for sc in resident]
    mix, batches = run(server, ids, list(range(len(requests))))
    assert max(batches.values()) > 1, "the mix must batch requests together"
    for i in range(len(requests)):
        for j, (a, b) in enumerate(zip(alone[i], mix[i])):
            assert not differ(a, b), f"request {i} (scene {requests[i][0]}) view {j} alone vs mix: {differ(a, b)}"
    server.close()
    # One request per scene against the oracle: its fewest-view request without a triangle budget.
    print("serving: oracle check of one request per scene", flush=True)
    checks = []
    for k in range(8):
        mine = [i for i, (kk, arr, n) in enumerate(requests) if kk == k and arr[0].triangle_budget == 0]
        assert mine, f"scene {k} has no request without a budget"
        checks.append(min(mine, key=lambda i: requests[i][2]))
    jobs = []
    for i in checks:
        k, arr, n = requests[i]
        pool = oracle.pool_from_scene(resident[k], mm, ctx.config.max_triangles, ctx.config.max_shadow_triangles)
        jobs += [(i, j, pool, arr[j]) for j in range(n)]

    def reference(job):
        i, j, pool, v = job
        ref = oracle.cut(pool, v)
        return i, j, oracle_compact(ref, mm), state_order_dependent(ref)

    with ThreadPoolExecutor(max_workers=16) as ex:
        for i, j, r, over in ex.map(reference, jobs):
            assert not over, f"request {i} view {j}: a round overflowed the draw list"
            assert not differ(alone[i][j], r), f"request {i} (scene {requests[i][0]}) view {j} vs oracle: {differ(alone[i][j], r)}"
    for sc in resident:
        sc.close()
    ctx.close()
