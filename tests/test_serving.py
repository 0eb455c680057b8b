"""The request server (vg_server_*): continuous batching over resident scenes.

Each request's result must be the same alone and in a mix of concurrent requests of several scenes
(bit-exact: counters, args and drawn sets), and equal to a plain cut of the same views from a fresh
controller state. Admission by HBM footprint and by free view slots is checked from the host side:
a budget that cannot hold a request refuses it, and a request that finds no free slots waits until
another request's slots are released.
"""
import time

import numpy as np
import pytest

from helpers import draw_set, ground_camera, small_ruins
from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenes
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu

MAX_MESHLETS = 1 << 18  # no round of any test view overflows the draw list (checked with the oracle)


def make_scenes(ctx, specs):
    out = []
    for spec in specs:
        scene = native.Scene(ctx)
        scenes.load_into(scene, spec)
        out.append(scene)
    return out


def request_views(rng, n):
    views = []
    for k in range(n):
        if rng.random() < 0.2:
            ortho = cam.orthographic_camera(-40, 40, 40, -40, 0.1, 400).look_at((rng.uniform(-20, 20), 150.0, rng.uniform(-20, 20)),
                                                                              (0.0, 0.0, 0.0))
            views.append(make_view(ortho, 1024, slot=k))
            continue
        eye = (rng.uniform(-50, 50), rng.uniform(2, 30), rng.uniform(-50, 50))
        target = (rng.uniform(-50, 50), 0.0, rng.uniform(-50, 50))
        budget = int(rng.integers(20_000, 200_000)) if rng.random() < 0.25 else 0
        views.append(make_view(ground_camera(eye, target), int(rng.choice([540, 720, 1080])), slot=k,
                               error_threshold=float(rng.uniform(0.5, 2.0)), triangle_budget=budget,
                               threshold_range=(0.5, 3.0) if budget else (0.5, 3.0)))
    return views


def results(server, ticket, count):
    out = []
    for i in range(count):
        counters, args, draws, _ = server.read(ticket, i, MAX_MESHLETS)
        sets = [draw_set(draws[c], min(int(counters[2 + 5 * c]), MAX_MESHLETS)) for c in (0, 1)]
        out.append((counters.copy(), args.copy(), sets))
    return out


def wait_all(server, tickets, timeout=60.0):
    done = {}
    t0 = time.time()
    while len(done) < len(tickets):
        for d in server.poll(256, wait_us=2000):
            done[d.ticket] = (d.status, d.batch_requests, d.lane)
        assert time.time() - t0 < timeout, f"requests did not complete: {len(done)} of {len(tickets)}"
    return done


def differences(a, b):
    """Where two results (lists of per-view (counters, args, sets)) differ, as text; empty when equal."""
    out = []
    for i, ((c1, a1, s1), (c2, a2, s2)) in enumerate(zip(a, b)):
        d = np.nonzero(c1 != c2)[0]
        if len(d):
            out.append(f"view {i} counters at {d[:8].tolist()}: {c1[d[:8]].tolist()} vs {c2[d[:8]].tolist()}")
        if not np.array_equal(a1, a2):
            out.append(f"view {i} args")
        for c in (0, 1):
            if not np.array_equal(s1[c], s2[c]):
                out.append(f"view {i} cut {c} sets ({len(s1[c])} vs {len(s2[c])})")
    return "; ".join(out)


def run_mix(server, ids, reqs, alone):
    """Submits every request at once and checks each against its alone result as it completes (then
    releases it, which frees view slots for the requests still queued). Returns the largest batch."""
    tickets = [server.submit(ids[sid], views) for sid, views in reqs]
    index = {t: k for k, t in enumerate(tickets)}
    pending, batched = set(tickets), 1
    t0 = time.time()
    while pending:
        for d in server.poll(256, wait_us=2000):
            k = index[d.ticket]
            assert d.status == native.VG_OK
            batched = max(batched, d.batch_requests)
            diff = differences(alone[k], results(server, d.ticket, len(reqs[k][1])))
            assert not diff, f"request {k} (scene {reqs[k][0]}, batch of {d.batch_requests}, lane {d.lane}) differs in the mix: {diff}"
            server.release(d.ticket)
            pending.discard(d.ticket)
        assert time.time() - t0 < 60, f"{len(pending)} requests did not complete"
    return batched


def test_each_request_is_the_same_alone_and_in_the_mix(gpu_available):
    specs = [small_ruins(seed=81, instances=800, meshes=3), small_ruins(seed=82, instances=2500, meshes=4),
             small_ruins(seed=83, instances=5000, meshes=5, shadows=True)]
    ctx = native.Context(max_triangles=2_000_000, max_shadow_triangles=1_000_000, max_meshlets=MAX_MESHLETS, max_views=16, device=0)
    resident = make_scenes(ctx, specs)
    server = native.Server(ctx, lanes=1, max_batch_views=16)
    ids = [server.add_scene(s) for s in resident]
    rng = np.random.default_rng(7)
    reqs = [(int(rng.integers(0, len(ids))), request_views(rng, int(rng.integers(1, 5)))) for _ in range(24)]
    # A draw list that overflows in some round makes the index counts depend on the append order (in the
    # reference too): such a view has no single answer, so the test views must not overflow.
    pools = [oracle.pool_from_scene(s, MAX_MESHLETS, 2_000_000, 1_000_000) for s in resident]
    for k, (sid, views) in enumerate(reqs):
        for v in views:
            r = oracle.cut(pools[sid], v)
            assert int(r.stats.draw_overflow[0]) + int(r.stats.draw_overflow[1]) == 0, f"request {k} overflows its draw list"

    # alone: one request at a time
    alone = []
    for sid, views in reqs:
        t = server.submit(ids[sid], views)
        status = wait_all(server, [t])[t][0]
        assert status == native.VG_OK
        alone.append(results(server, t, len(views)))
        server.release(t)

    # in the mix: everything at once, so requests are batched (one lane: no overlap; three lanes: batches
    # overlap on the device)
    batched = run_mix(server, ids, reqs, alone)
    assert batched > 1, "no two requests were batched together"
    server.close()
    server = native.Server(ctx, lanes=3, max_batch_views=16)
    ids = [server.add_scene(s) for s in resident]
    run_mix(server, ids, reqs, alone)
    server.close()
    server = native.Server(ctx, lanes=2, max_batch_views=16, policy=native.VG_SERVE_FIFO)
    ids = [server.add_scene(s) for s in resident]
    run_mix(server, ids, reqs, alone)
    server.close()
    # launches capped at 2 persistent blocks per CU (room for the other lanes' kernels): same results
    server = native.Server(ctx, lanes=3, max_batch_views=16, max_blocks_per_cu=2)
    ids = [server.add_scene(s) for s in resident]
    run_mix(server, ids, reqs, alone)
    server.close()
    # a first work list far too small: every batch overflows it and runs again at full size
    server = native.Server(ctx, lanes=2, max_batch_views=16, first_work_items=64)
    ids = [server.add_scene(s) for s in resident]
    run_mix(server, ids, reqs, alone)
    assert server.stats().reruns > 0
    st = server.stats()
    assert st.batches - st.reruns < len(reqs) and st.requests_done == len(reqs)
    assert st.resident_bytes + st.peak_inflight_bytes <= st.hbm_budget

    # alone == a plain cut of the same views from a fresh controller (the plain path on its own slots)
    for k in (0, 5, 11):
        sid, views = reqs[k]
        scene = resident[sid]
        for v in views:
            scene.set_state(v.slot, oracle.initial_state())
        scene.cut(views)
        scene.sync()
        for i, v in enumerate(views):
            counters, args, draws, _ = scene.read(v.slot, MAX_MESHLETS)
            sets = [draw_set(draws[c], min(int(counters[2 + 5 * c]), MAX_MESHLETS)) for c in (0, 1)]
            diff = differences([alone[k][i]], [(counters, args, sets)])
            assert not diff, f"request {k} view {i}: served != plain: {diff}"
    server.close()


def test_admission_by_hbm_budget_and_view_slots(gpu_available):
    specs = [small_ruins(seed=91, instances=1000, meshes=3), small_ruins(seed=92, instances=4000, meshes=4)]
    ctx = native.Context(max_triangles=2_000_000, max_shadow_triangles=1_000_000, max_meshlets=MAX_MESHLETS, max_views=4, device=0)
    small, large = make_scenes(ctx, specs)
    # A budget that holds both scenes and two small 2-view requests (about 0.46 MB of cell queue and work
    # list each: 16 work-list partitions of up to 4 queue entries x 256 items) but not a 4-view request of
    # the large scene (about 1.26 MB: partitions of up to 11 entries).
    probe = native.Server(ctx, lanes=1)
    probe.add_scene(small)
    probe.add_scene(large)
    resident = probe.stats().resident_bytes
    probe.close()
    server = native.Server(ctx, hbm_budget=resident + 1_000_000, lanes=2, max_batch_views=4)
    s_id, l_id = server.add_scene(small), server.add_scene(large)
    rng = np.random.default_rng(3)
    with pytest.raises(native.VgError) as e:
        server.submit(l_id, request_views(rng, 4))
    assert e.value.code == native.VG_E_NOMEM
    # View slots: 4 per scene. Two 2-view requests take them; a third waits for a release.
    t1 = server.submit(s_id, request_views(rng, 2))
    t2 = server.submit(s_id, request_views(rng, 2))
    wait_all(server, [t1, t2])
    t3 = server.submit(s_id, request_views(rng, 1))
    assert not server.poll(16, wait_us=200_000), "a request ran without free view slots"
    server.release(t1)
    done = wait_all(server, [t3])
    assert done[t3][0] == native.VG_OK
    server.release(t2)
    server.release(t3)
    server.close()


def test_shortest_first_and_fifo_order(gpu_available):
    """One lane: while a first request runs, a 16-view request of one scene and then a 1-view request of
    another queue behind it. Shortest-first launches the 1-view request first; FIFO keeps arrival order."""
    specs = [small_ruins(seed=95, instances=800, meshes=3), small_ruins(seed=96, instances=2500, meshes=3),
             small_ruins(seed=97, instances=1500, meshes=3)]
    ctx = native.Context(max_triangles=2_000_000, max_shadow_triangles=1_000_000, max_meshlets=MAX_MESHLETS, max_views=16, device=0)
    resident = make_scenes(ctx, specs)
    rng = np.random.default_rng(5)
    first, heavy, light = request_views(rng, 4), request_views(rng, 16), request_views(rng, 1)
    for policy, light_first in ((native.VG_SERVE_SHORTEST, True), (native.VG_SERVE_FIFO, False)):
        server = native.Server(ctx, lanes=1, max_batch_views=16, policy=policy)
        ids = [server.add_scene(s) for s in resident]
        ta = server.submit(ids[2], first)
        tb = server.submit(ids[1], heavy)
        tc = server.submit(ids[0], light)
        queued = {}
        t0 = time.time()
        while len(queued) < 3:
            for d in server.poll(8, wait_us=2000):
                assert d.status == native.VG_OK
                queued[d.ticket] = d.queued_us
                server.release(d.ticket)
            assert time.time() - t0 < 30
        assert (queued[tc] < queued[tb]) == light_first, (policy, queued[ta], queued[tb], queued[tc])
        server.close()


def test_full_size_reruns_wait_for_hbm_budget(gpu_available):
    """A batch whose optimistic work list overflows runs again at full size only when that fits the HBM
    budget: two overflowing requests whose full-size transients do not fit together run one after the
    other (same results as alone), and one whose full size cannot fit even alone completes with
    VG_E_NOMEM instead of failing an allocation."""
    specs = [small_ruins(seed=97, instances=2500, meshes=3), small_ruins(seed=98, instances=2500, meshes=3)]
    ctx = native.Context(max_triangles=2_000_000, max_shadow_triangles=1_000_000, max_meshlets=MAX_MESHLETS, max_views=4, device=0)
    resident = make_scenes(ctx, specs)
    rng = np.random.default_rng(9)
    reqs = [request_views(rng, 4), request_views(rng, 4)]
    # Alone, with room to spare: each request's full-size transient bytes (the peak after its rerun).
    alone, full = [], []
    for k in range(2):
        probe = native.Server(ctx, lanes=1, max_batch_views=4, first_work_items=64)
        ids = [probe.add_scene(s) for s in resident]
        t = probe.submit(ids[k], reqs[k])
        assert wait_all(probe, [t])[t][0] == native.VG_OK
        alone.append(results(probe, t, 4))
        st = probe.stats()
        assert st.reruns == 1
        full.append(int(st.peak_inflight_bytes))
        base = int(st.resident_bytes)
        probe.close()
    # Room for either full-size rerun, not both.
    budget = base + max(full) + min(full) // 2
    server = native.Server(ctx, hbm_budget=budget, lanes=2, max_batch_views=4, first_work_items=64)
    ids = [server.add_scene(s) for s in resident]
    tickets = [server.submit(ids[k], reqs[k]) for k in range(2)]
    done = wait_all(server, tickets)
    for k, t in enumerate(tickets):
        assert done[t][0] == native.VG_OK
        diff = differences(alone[k], results(server, t, 4))
        assert not diff, f"request {k} after a deferred rerun: {diff}"
        server.release(t)
    st = server.stats()
    assert st.reruns == 2 and st.resident_bytes + st.peak_inflight_bytes <= budget
    server.close()
    # No room for the full size at all.
    server = native.Server(ctx, hbm_budget=base + min(full) - 1, lanes=1, max_batch_views=4, first_work_items=64)
    ids = [server.add_scene(s) for s in resident]
    k = int(np.argmin(full))
    t = server.submit(ids[k], reqs[k])
    assert wait_all(server, [t])[t][0] == native.VG_E_NOMEM
    server.release(t)
    server.close()
