#!/usr/bin/env python3
"""Cut benchmark on the 3M-instance world: views x clusters tested per second, plain vs batched.

Workload: the world scene (64 procedural meshes of 0.1M-2M triangles, 3.0M instances by Poisson-disk
sampling over a 6144 m x 6144 m heightfield, per-mesh draw distances 200-6000 m) cut for 64 mixed
views: 4 main cameras (2 at 1920x1080, 2 at 3840x2160, device pixels), 16 orthographic shadow
cascades at 2048^2, 24 cube-map faces at 512^2 (90 degrees, 60 m far) and 20 viewers at 1280x720.
Perspective views with at least 720 lines are occlusion culled against the rasterized terrain.

A step cuts all 64 views: the plain form launches one view per launch sequence (64 launches), the
batched form cuts them in one launch sequence replayed from a HIP graph. `value` is the batched
form's views x clusters tested per second with the scene resident in HBM.

Beside it: the expand pass (index writing) on the same views at the reference's default capacities,
and served request streams over 8 resident scenes (batched server vs one request per launch).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "three-virtual-geometry_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Draw capacity per view: the reference's defaults are 10M / 4M triangles per pool (VirtualGeometry.ts:191-198);
# the world's main views ask for up to ~60M triangles at 1 px, so the buffers are sized for MI355X instead
# (no capacity retries at 1 px; a 300M-triangle index buffer is 3.6 GB).
MAX_TRIANGLES = 300_000_000
MAX_SHADOW_TRIANGLES = 100_000_000
# The reference's default pool capacities (VirtualGeometry.ts:191-198, GeometryPool.ts:642-644): the expand
# measurement runs at these, so the 64 views' index buffers (triangles x 3 u32 per cut) take 10.75 GB.
REF_TRIANGLES = 10_000_000
REF_SHADOW_TRIANGLES = 4_000_000
REF_MESHLETS = 1 << 20


TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "world64_rocprof.json")
EXPAND_TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "world64_expand_rocprof.json")
RASTER_TRAFFIC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "world24_raster_rocprof.json")


def measured_traffic(kernel: str, path: str = TRAFFIC_SUMMARY):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 summary of this workload (two --pmc
    runs, FETCH_SIZE and WRITE_SIZE, calibrated on a device copy), or (None, reason)."""
    rel = os.path.relpath(path, os.path.dirname(os.path.abspath(__file__)))
    if not os.path.exists(path):
        return None, f"no rocprofv3 summary ({rel})"
    with open(path) as f:
        summary = json.load(f)
    for name, k in summary.get("kernels", {}).items():
        if name.split("<")[0] == kernel and "hbm_bytes" in k:
            return k["hbm_bytes"], (f"{rel}: {name}, FETCH_SIZE x {summary['fetch_factor']:.2f} + "
                                    f"WRITE_SIZE x {summary['write_factor']:.2f} per dispatch, avg dispatch {k['avg_us']:.0f} us")
    return None, f"{kernel} not in the rocprofv3 summary"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--instances", type=int, default=3_000_000)
    ap.add_argument("--meshes", type=int, default=64)
    ap.add_argument("--views", type=int, default=64)
    ap.add_argument("--latency-samples", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound on the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="a few batched launches for rocprofv3")
    ap.add_argument("--check", action="store_true", help="compare the batched cut of every view with the oracle")
    ap.add_argument("--max-meshlets", type=int, default=1 << 22, help="draw-list entries per cut per view")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-requests measurement")
    ap.add_argument("--serve-only", action="store_true", help="only the served-requests measurement")
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="requests per rate = rate x this (at least 1000)")
    ap.add_argument("--no-expand", action="store_true", help="skip the expand-pass measurement")
    ap.add_argument("--plain-configs", action="store_true", help="only the plain single-view configurations (icosphere, ruins)")
    ap.add_argument("--expand-only", action="store_true", help="only the expand-pass measurement (with --profile-only: for rocprofv3)")
    ap.add_argument("--two-phase-only", action="store_true", help="only the two-phase occlusion frames with their rasters (for rocprofv3)")
    ap.add_argument("--schedule", type=str, default=None,
                    help="fixed launch schedule 'instance_blocks,meshlet_blocks,meshlet_waves' (per CU) instead of tuning")
    return ap.parse_args()


SERVE_RATES = (2_000, 5_000, 10_000, 20_000)


def serve_stream(server, ids, pool, rate, count, seed):
    """Poisson arrivals at `rate` req/s of `count` requests drawn in order from the pool; completed
    requests are released as they are seen. Returns (offered req/s, elapsed s, done records)."""
    import ctypes

    from vgcut import native

    lib, handle = server.lib, server.handle
    done_buf = (native.RequestDone * 512)()
    n_done = ctypes.c_uint32()
    ticket = ctypes.c_uint64()
    gaps = np.random.default_rng(seed).exponential(1.0 / rate, count)
    records = []
    which = {}

    def drain(wait_us):
        lib.vg_server_poll(handle, done_buf, 512, ctypes.byref(n_done), wait_us)
        for j in range(n_done.value):
            d = done_buf[j]
            records.append((which.pop(d.ticket), d.status, d.batch_requests, d.total_us, d.queued_us))
            lib.vg_server_release(handle, d.ticket)

    t0 = time.perf_counter()
    t_next = t0
    for i in range(count):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            drain(0)
        k, arr, n = pool[i % len(pool)]
        rc = lib.vg_server_submit(handle, ids[k], ctypes.cast(arr, ctypes.c_void_p), n, ctypes.byref(ticket))
        if rc != 0:
            raise RuntimeError(f"vg_server_submit failed ({rc})")
        which[ticket.value] = i % len(pool)
    t_sub = time.perf_counter()
    while len(records) < count:
        drain(1000)
    return count / (t_sub - t0), time.perf_counter() - t0, records


def serve_bench(spec_meshes, make_view, cam, scenegen, rates, requests_per_rate, max_meshlets):
    """Served cut requests over 8 resident scenes (12k-3M instances): batched server vs one request per
    launch, Poisson arrivals at each rate. Each view slot has the draw capacity of the world's views."""
    from vgcut import native, scenes

    t0 = time.time()
    ctx = native.Context(max_triangles=MAX_TRIANGLES, max_shadow_triangles=MAX_SHADOW_TRIANGLES, max_meshlets=max_meshlets, max_views=48,
                         mesh_slots=64, device=0)
    specs = scenes.serving_specs(spec_meshes)
    resident = []
    for spec, side in specs:
        scene = native.Scene(ctx)
        scenes.load_into(scene, spec)
        resident.append((scene, spec, side))
    pool = scenes.serve_requests(specs, 256)
    # Each scene's launch schedule, tuned once on its largest request of the pool (both servers use it).
    for k, (scene, _, _) in enumerate(resident):
        mine = [(n, arr) for kk, arr, n in pool if kk == k]
        if mine:
            n, arr = max(mine, key=lambda r: r[0])
            scene.tune(list(arr)[:n], repeats=3)
    log(f"serving: {len(scenes.SERVE_SIZES)} scenes, {sum(s.instance_count for _, s, _ in resident)} instances, "
        f"schedules tuned ({time.time() - t0:.1f} s)")

    def server_for(batched, policy=native.VG_SERVE_SHORTEST):
        # Batched: 4 lanes whose launches take at most 4 persistent blocks per CU, leaving room for the other
        # lanes' kernels (light requests then wait less behind heavy ones). Plain: one request per launch.
        sv = native.Server(ctx, lanes=4 if batched else 1, max_batch_views=48, max_batch_requests=0 if batched else 1,
                           policy=policy if batched else native.VG_SERVE_FIFO, max_blocks_per_cu=4 if batched else 0)
        return sv, [sv.add_scene(sc) for sc, _, _ in resident]

    # request classes: views (light 1-4, heavy 5-16) x scene (small <= 152k instances, large >= 500k)
    def klass(k):
        _, sid, n = k
        return ("light" if n <= 4 else "heavy") + ("_small" if scenes.SERVE_SIZES[sid] <= 152_000 else "_large")

    classes = [klass((None, k, n)) for k, _, n in pool]

    # Work of each pool request (cluster tests) and the plain service time, from one pass alone.
    plain, plain_ids = server_for(False)
    tests, service = [], []
    for k, arr, n in pool:
        t = plain.submit_array(plain_ids[k], arr, n)
        d = []
        while not d:
            d = plain.poll(4, wait_us=100_000)
        service.append(d[0].run_us)
        mets = [plain.read(t, i, 1)[3] for i in range(n)]
        tests.append(sum(int(m.clusters_tested[0] + m.clusters_tested[1]) for m in mets))
        plain.release(t)
    capacity = 1e6 / float(np.mean(service))  # plain requests per second, alone and back to back
    out = {"scenes_instances": list(scenes.SERVE_SIZES), "request_pool": len(pool), "views_per_request_mean": float(np.mean([p[2] for p in pool])),
           "cluster_tests_per_request_mean": float(np.mean(tests)), "plain_service_ms_mean": float(np.mean(service)) / 1e3,
           "plain_capacity_req_per_s": capacity, "hbm_resident_bytes": int(plain.stats().resident_bytes), "rates": []}
    out["class_counts"] = {c: classes.count(c) for c in sorted(set(classes))}
    out["class_plain_service_ms_mean"] = {c: float(np.mean([service[i] for i in range(len(pool)) if classes[i] == c])) / 1e3
                                          for c in sorted(set(classes))}
    batched, batched_ids = server_for(True)
    fifo, fifo_ids = server_for(True, native.VG_SERVE_FIFO)
    legs = (("batched", batched, batched_ids), ("batched_fifo", fifo, fifo_ids), ("plain", plain, plain_ids))
    # Warm-up: every pool request once through each server, so its lanes hold the buffers the mix needs and
    # its per-scene run-time estimates exist (a server in service is past this; the plain one is already).
    for name, sv, ids in legs[:2]:
        serve_stream(sv, ids, pool, capacity, len(pool), seed=0)
    loads = sorted({int(capacity * f) for f in (0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0)} | set(rates))
    for rate in loads:
        # about `requests_per_rate` seconds of arrivals, and at most about 2 s of work at plain capacity
        count = int(max(100, min(rate * requests_per_rate, 2.0 * capacity)))
        row = {"offered_req_per_s": rate, "requests": count}
        for name, sv, ids in legs:
            offered, elapsed, rec = serve_stream(sv, ids, pool, rate, count, seed=rate)
            lat = [r[3] / 1e3 for r in rec]
            ok = sum(1 for r in rec if r[1] == 0)
            row[name] = {"submitted_req_per_s": offered, "completed_req_per_s": count / elapsed,
                         "views_clusters_per_s": sum(tests[r[0]] for r in rec) / elapsed,
                         "latency_ms_p50": percentile(lat, 50), "latency_ms_p99": percentile(lat, 99),
                         "mean_batch_requests": float(np.mean([r[2] for r in rec])), "ok": ok, "classes": {}}
            for c in sorted(set(classes)):
                cl = [r[3] / 1e3 for r in rec if classes[r[0]] == c]
                if cl:
                    row[name]["classes"][c] = {"n": len(cl), "p50_ms": percentile(cl, 50), "p99_ms": percentile(cl, 99)}
            light = [r[3] / 1e3 for r in rec if classes[r[0]].startswith("light")]
            row[name]["light_p99_ms"] = percentile(light, 99) if light else None
        for name, sv, _ in legs:
            st = sv.stats()
            row[name]["peak_inflight_bytes"] = int(st.peak_inflight_bytes)
            row[name]["reruns_so_far"] = int(st.reruns)
        out["rates"].append(row)
        log(f"serving {rate} req/s: " + "; ".join(
            f"{name} p50 {row[name]['latency_ms_p50']:.2f} p99 {row[name]['latency_ms_p99']:.2f} ms light p99 {row[name]['light_p99_ms']:.2f} "
            f"({row[name]['completed_req_per_s']:.0f} req/s, batch {row[name]['mean_batch_requests']:.1f})" for name, _, _ in legs))
    for name, _, _ in legs:
        ok_rates = [r["offered_req_per_s"] for r in out["rates"] if r[name]["light_p99_ms"] is not None and r[name]["light_p99_ms"] < 2.0]
        out[f"{name}_light_p99_under_2ms_max_offered_req_per_s"] = max(ok_rates) if ok_rates else 0
        out[f"{name}_capacity_req_per_s"] = max(r[name]["completed_req_per_s"] for r in out["rates"])
    batched.close()
    fifo.close()
    plain.close()
    return out


def expand_bench(spec, vd, steps, warmup, schedule, profile_only=False):
    """The cut followed by the expand pass (GeometryPool.ts:997-1021) for the 64 views at the reference's
    default pool capacities (10M camera / 4M shadow triangles, 1M meshlets per view): the main views'
    demand exceeds them, so the capacity controller retries those cuts coarser, as in the reference. The
    expand pass then writes 3 indices per drawn triangle into each view's index buffers."""
    import torch

    from vgcut import native, scenes

    views = [v for v, _ in vd]
    ctx = native.Context(max_triangles=REF_TRIANGLES, max_shadow_triangles=REF_SHADOW_TRIANGLES, max_meshlets=REF_MESHLETS,
                         max_views=len(views), mesh_slots=72, device=0, expand=True)
    scene = native.Scene(ctx)
    t0 = time.time()
    scenes.load_into(scene, spec, occlusion=True)
    for v, d in vd:
        if d is not None:
            scene.set_hzb(v.slot, d)
    scene.set_schedule(*schedule)
    for _ in range(max(1, warmup)):
        scene.cut(views, graph=True)
    scene.sync()
    log(f"expand: scene loaded at the reference capacities, warm ({time.time() - t0:.1f} s)")
    if profile_only:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        for _ in range(steps):
            scene.cut(views, graph=False)
        scene.sync()
        return {"profile_only": True, "steps": steps, "views": len(views), "copy_bytes": 2 * src.numel() * 4}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        scene.cut(views, graph=True)
    scene.sync()
    step_s = (time.perf_counter() - t0) / steps
    scene.profile(True)
    for _ in range(max(3, steps // 2)):
        scene.cut(views, graph=False)
    scene.sync()
    kernels, seqs = scene.profile_read()
    scene.profile(False)
    kernel_ms = {name: ms / max(1, seqs) for name, ms in kernels}
    expand_ms = kernel_ms["k_expand (+ plan, resolve)"]

    # The last cut's work: drawn entries, indices written (exact, also where a region passes the capacity)
    # and the triangles of the distinct meshlets drawn (their triangle words are read at least once).
    slots = ctx.config.mesh_slots
    info = scene.pool(native.VG_POOL_MESHLET_INFO).reshape(-1, 4)
    slot_of = scene.pool(native.VG_POOL_INSTANCES).reshape(-1, 20)[:, 19].astype(np.int64)
    drawn = written = 0
    drawn_meshlets = []
    for v in views:
        counters = scene.read(v.slot, max_entries=1)[0]
        n = (int(counters[3]), int(counters[8]))
        draws = scene.read(v.slot, max_entries=max(1, max(n)))[2]
        for cut in (0, 1):
            d = draws[cut, :n[cut]].astype(np.int64)
            drawn += n[cut]
            drawn_meshlets.append(d[:, 1])
            requested = int(counters[2 + cut * 5 + 4])
            cap = (REF_TRIANGLES if cut == 0 else REF_SHADOW_TRIANGLES) * 3
            if requested <= cap:
                written += requested
            else:
                part = np.where((d[:, 3] & 0xFFFF00) == (255 << 16), 0, 1)
                region = counters[12 + cut * slots * 4 + 1 + slot_of[d[:, 0]] * 4 + part * 2].astype(np.int64)
                first = region + d[:, 2]
                fits = np.clip((cap - 2 - first + 2) // 3, 0, d[:, 3] & 255)  # triangles i with first + 3i + 2 < cap
                written += 3 * int(fits.sum())
    unique = np.unique(np.concatenate(drawn_meshlets)) if drawn_meshlets else np.zeros(0, dtype=np.int64)
    unique_triangles = int(info[unique, 0].astype(np.int64).sum())
    algorithmic = 24 * drawn + 4 * written + 4 * unique_triangles
    achieved = algorithmic / (expand_ms * 1e-3) / 1e9
    traffic, source = 0.0, []
    for name in ("k_expand_plan", "k_expand_resolve", "k_expand"):
        t, src = measured_traffic(name, EXPAND_TRAFFIC_SUMMARY)
        if t is None:
            traffic, source = None, [src]
            break
        traffic += t
        source.append(src)
    source = "; ".join(source)
    out = {"workload": f"world {spec.instance_count} instances, {len(views)} mixed views, capacities 10M/4M triangles, 1M meshlets",
           "ms_per_step_cut_and_expand": step_s * 1e3, "kernel_ms": kernel_ms, "expand_ms": expand_ms,
           "drawn_meshlets": drawn, "indices_written": written, "distinct_meshlets": int(len(unique)),
           "indices_per_s": written / (expand_ms * 1e-3),
           "roofline": {"bound": "hbm", "kernel": "k_expand_plan + k_expand_resolve + k_expand", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": algorithmic, "traffic": traffic,
                        "traffic_source": source}}
    log(f"expand: {drawn} drawn meshlets, {written} indices in {expand_ms:.3f} ms ({achieved:.0f} GB/s algorithmic), "
        f"cut + expand {step_s * 1e3:.2f} ms per step")
    scene.close()
    ctx.close()
    return out


def two_phase_bench(scene, vd, steps):
    """Two-phase occlusion against the previous frame's visible set (config 2's occlusion, OcclusionCulling.ts:4-21):
    on each view with an HZB (occlusion = 2), the occluder raster draws the slot's last cut, both phases, with
    the camera into its HZB (vg_hzb_from_cut, on the device); the cut records its camera occlusion rejects;
    the raster draws that cut's clusters; the second phase re-tests the rejects against them. Reports each
    step's time, the rasters' share, what was re-tested and added, and the raster kernels' roofline (their
    event-timed duration, algorithmic bytes, HBM traffic from the committed rocprof summary)."""
    import ctypes

    import torch

    from vgcut import camera as cam
    from vgcut import native

    # The occlusion-culled views move: each frame every one of them steps `speed` metres along its own forward
    # axis (the view matrix's z translation), so the previous frame's visible set, drawn from the new camera,
    # leaves gaps the second phase fills. The other views stay where they are.
    speed = 1.5
    occluded = [bool(d is not None and v.occlusion) for v, d in vd]
    size = {v.slot: (d.shape[1], d.shape[0]) for v, d in vd if d is not None}

    def views_at(f):
        views, rasters = [], []
        for (v, _), occl in zip(vd, occluded):
            w = native.View()
            ctypes.memmove(ctypes.byref(w), ctypes.byref(v), ctypes.sizeof(native.View))
            if occl:
                V = np.array(v.view_matrix[:], dtype=np.float64)
                V[14] += speed * f
                P = np.array(v.projection_matrix[:], dtype=np.float64)
                planes = cam.frustum_planes(cam.mat_mul(P, V)).reshape(24)
                for i in range(16):
                    w.view_matrix[i] = float(V[i])
                for i in range(24):
                    w.frustum_planes[i] = float(planes[i])
                w.occlusion = 2
                vp = np.asarray(cam.mat_mul(P, np.array(w.view_matrix[:], dtype=np.float64)), dtype=np.float32)
                rasters.append((w.slot, vp) + size[w.slot])
            views.append(w)
        return views, [v for v in views if v.occlusion == 2], rasters

    warm = 3
    frames = [views_at(f) for f in range(warm + steps)]

    def frame(f):
        views, two, rasters = frames[f]
        t = [0.0, 0.0, 0.0, 0.0]
        def mark(i, t0):
            scene.sync()
            t[i] += time.perf_counter() - t0
        t0 = time.perf_counter()
        if f > 0:  # every occluded view's raster in one launch sequence (vg_hzb_from_cuts)
            scene.hzb_from_cuts([(slot, vp, w, h, True) for slot, vp, w, h in rasters])
        mark(0, t0)
        t0 = time.perf_counter()
        scene.cut(views, graph=True)
        mark(1, t0)
        t0 = time.perf_counter()
        scene.hzb_from_cuts([(slot, vp, w, h, False) for slot, vp, w, h in rasters])
        mark(2, t0)
        t0 = time.perf_counter()
        scene.cut_phase2(two)
        mark(3, t0)
        return t

    for f in range(warm):
        frame(f)
    total = [0.0, 0.0, 0.0, 0.0]
    added_frames = []
    scene.profile_read()
    scene.profile(True)
    for f in range(warm, warm + steps):
        torch.cuda.synchronize()
        total = [a + b for a, b in zip(total, frame(f))]
        added_frames.append(sum(int(scene.read_phase2(v.slot)[0][2]) - int(scene.read(v.slot, max_entries=1)[0][2])
                                for v in frames[f][1]))
    views, two, rasters = frames[-1]
    kernels, _ = scene.profile_read()
    scene.profile(False)
    raster_kernel_ms = dict(kernels)["k_raster (+ chunks)"] / steps  # both rasters of a frame
    ms = [x / steps * 1e3 for x in total]
    rejected_meshlets = sum(scene.phase_rejects(v.slot)[0] for v in two)
    rejected_instances = added = drawn = 0
    raster_bytes = 0
    pool_info = scene.pool(native.VG_POOL_MESHLET_INFO).reshape(-1, 4)
    pool_tri = scene.pool(native.VG_POOL_MESHLET_TRIANGLES)

    def entry_bytes(meshlets):
        # Per drawn camera entry: the entry, 4 matrix columns, meshlet records (first, count, vertex list),
        # the positions (16 B, stored in vertex-list order), the packed triangles (vertex count = largest
        # local id + 1).
        info = pool_info[meshlets]
        has = info[:, 3] != 0xFFFFFFFF
        tris = info[:, 0].astype(np.int64)
        nv = np.zeros(len(meshlets), dtype=np.int64)
        for k in np.nonzero(has)[0]:
            p = pool_tri[info[k, 1] // 3: info[k, 1] // 3 + tris[k]]
            nv[k] = int(max((p & 255).max(), ((p >> 8) & 255).max(), ((p >> 16) & 255).max())) + 1
        return int(len(meshlets) * (16 + 64 + 16) + (nv * 16).sum() + (tris[has] * 4).sum())

    for v in two:
        c1, _, draws, _ = scene.read(v.slot)
        c2, _ = scene.read_phase2(v.slot)
        rejected_instances += int(c1[1])
        added += int(c2[2]) - int(c1[2])
        mm = scene.ctx.config.max_meshlets
        n1, n2 = min(int(c1[2]), mm), min(int(c2[2]), mm)
        drawn += n1
        # A frame's two rasters: the previous frame's cut (both phases; counted as this frame's, which the
        # camera's small step leaves about the same) and this frame's first phase.
        raster_bytes += entry_bytes(draws[0, :n2, 1].astype(np.int64)) + entry_bytes(draws[0, :n1, 1].astype(np.int64))
    achieved = raster_bytes / (raster_kernel_ms * 1e-3) / 1e9 if raster_kernel_ms > 0 else None
    traffic, source = 0.0, []
    for name in ("k_raster", "k_raster_chunks"):
        t, src = measured_traffic(name, RASTER_TRAFFIC_SUMMARY)
        if t is None:
            traffic, source = None, [src]
            break
        traffic += t * 2  # per dispatch; one launch sequence per phase draws every view's raster, two per frame
        source.append(src)
    out = {"views_two_phase": len(two), "camera_step_m_per_frame": speed, "ms_per_frame": sum(ms), "raster_previous_ms": ms[0],
           "cut_ms_with_recording": ms[1], "raster_first_phase_ms": ms[2], "phase2_ms": ms[3], "instances_rejected": rejected_instances,
           "meshlets_rejected": rejected_meshlets, "phase2_clusters_added": added, "phase2_clusters_added_per_frame": added_frames,
           "raster_entries_first_phase": drawn,
           "raster": {"kernel_ms_per_frame": raster_kernel_ms, "rasters_per_frame": 2 * len(rasters),
                      "roofline": {"bound": "hbm", "kernel": "k_raster + k_raster_chunks", "achieved": achieved, "peak": HBM_PEAK_GBS,
                                   "unit": "GB/s", "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
                                   "algorithmic_bytes": raster_bytes, "traffic": traffic, "traffic_source": "; ".join(source)}}}
    log(f"two-phase: {out['ms_per_frame']:.2f} ms per frame: raster of the last frame {ms[0]:.2f}, cut {ms[1]:.2f}, raster of "
        f"its clusters {ms[2]:.2f} ({drawn} entries), second phase {ms[3]:.2f}; raster kernels {raster_kernel_ms:.2f} ms per frame, "
        f"{raster_bytes / 1e9:.2f} GB algorithmic ({'-' if achieved is None else f'{achieved:.0f}'} GB/s); "
        f"{rejected_instances} instances and {rejected_meshlets} meshlets re-tested, {added} clusters added")
    initial = np.zeros(8, dtype=np.uint32)
    initial[0] = initial[1] = 10000
    for v in views:
        scene.set_state(v.slot, initial)
    return out


def plain_configs_bench(cpu_seconds):
    """The plain single-view configurations, each view cut by its own launch sequence in the controller's
    steady state: the 1M-triangle icosphere (one 1080p view at 1 px), and the 113k-instance ruins (8 poses at
    1080p and 2160p, 2 m eye height, the heightfield's depth as HZB), at the reference's default capacities
    (10M / 4M triangles) with an 8M-entry draw list, graph replay; per view: launch-to-completion latency
    p50/p99 over repeated launches, views*clusters tested/s and visible clusters/s; for the icosphere also the
    Node CPU cut (one thread), the reference's CPU-runnable case. Then the world's fly-through (below)."""
    import math

    import torch

    from oracle import oracle
    from vgcut import camera as cam
    from vgcut import native, scenegen, scenes
    from vgcut.virtual_geometry import make_view

    def run(scene, views, repeats):
        lat, tests, drawn = [], 0, 0
        for v in views:
            for _ in range(6):  # uploads, graph capture, controller steady state
                scene.cut([v], graph=True)
            scene.sync()
            counters, _, _, met = scene.read(v.slot, max_entries=1)
            tests += int(met.clusters_tested[0] + met.clusters_tested[1])
            drawn += int(counters[2]) + int(counters[7])
            for _ in range(repeats):
                t = time.perf_counter()
                scene.cut([v], graph=True)
                scene.sync()
                lat.append((time.perf_counter() - t) * 1e3)
        per_view_s = float(np.mean(lat)) / 1e3
        return {"views": len(views), "cluster_tests_per_view": tests / len(views), "drawn_per_view": drawn / len(views),
                "latency_ms_p50": percentile(lat, 50), "latency_ms_p99": percentile(lat, 99),
                "views_clusters_per_s": tests / len(views) / per_view_s, "visible_clusters_per_s": drawn / len(views) / per_view_s}

    out = {}
    torch.zeros(1, device="cuda:0")
    ctx = native.Context(max_meshlets=8 << 20, max_views=16, device=0)
    t0 = time.time()
    spec, camera, height = scenes.icosphere_view()
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec)
    view = make_view(camera, int(height), slot=0)
    out["config0_icosphere"] = run(scene, [view], 200)
    pool = oracle.pool_from_scene(scene, ctx.config.max_meshlets, ctx.config.max_triangles, ctx.config.max_shadow_triangles)
    scale = scene.state(0)[0] / 10000.0
    node = oracle.node_cpu_baseline(pool, [view], [scale], seconds=cpu_seconds, threads=1)
    one = node["single_thread"]
    out["config0_icosphere"]["cpu_baseline"] = {"value": one["tests_per_s"], "unit": "views*clusters tested/s", "cores": 1, "kind": "port",
                                               "cpu_model": node["cpu_model"], "runtime": f"node {node['node']}",
                                               "sample": f"oracle/cpu_cut.js, one thread: {one['tests']} cluster tests in {one['seconds']:.2f} s"}
    log(f"config 0 ({time.time() - t0:.1f} s): {out['config0_icosphere']}")
    scene.close()
    del pool

    t0 = time.time()
    spec = scenes.ruins()
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=True)
    hf = spec.heightfield
    rng = np.random.default_rng(4)
    views = []
    for pose in range(8):
        eye = (float(rng.uniform(-400, 400)), 2.0, float(rng.uniform(-400, 400)))
        yaw = float(rng.uniform(0, 2 * math.pi))
        target = (eye[0] + 10 * math.cos(yaw), 2.0, eye[2] + 10 * math.sin(yaw))
        for k, (h, hw, hh) in enumerate(((1080, 960, 540), (2160, 1024, 576))):
            camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at(eye, target)
            vp = cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse)
            scene.set_hzb(pose * 2 + k, scenegen.heightfield_depth(vp, hw, hh, (0.0, 0.0), 500.0, 256, hf["amplitude"], hf["frequency"], hf["seed"]))
            views.append(make_view(camera, h, slot=pose * 2 + k, occlusion=True))
    out["config1_ruins"] = {"1080p": run(scene, views[0::2], 40), "2160p": run(scene, views[1::2], 40)}
    log(f"config 1 ({time.time() - t0:.1f} s): {out['config1_ruins']}")
    scene.close()

    # The world's 16-frame fly-through at 1080p with the 8M-triangle budget (GPU histogram pick) and two-phase
    # occlusion: the first phase against the raster of the previous frame's visible set, the second against
    # the raster of the first phase's clusters (960x540 depth).
    t0 = time.time()
    spec = scenes.world(terrain_grid=512)
    wctx = native.Context(max_triangles=20_000_000, max_shadow_triangles=10_000_000, max_meshlets=4 << 20, max_views=1,
                          mesh_slots=72, device=0)
    scene = native.Scene(wctx)
    scenes.load_into(scene, spec, occlusion=True)
    path = scenes.flythrough(spec, 16, 6144.0, 30.0, seed=9)
    w, h = 960, 540
    frames = []
    for lap in range(2):  # the first lap warms up (uploads, graph capture, the controller)
        scene.set_hzb(0, np.ones((h, w), dtype=np.float32))
        for f in range(16):
            eye, target = path[f]
            camera = cam.perspective_camera(60, 16 / 9, 0.1, 20000).look_at(eye, target)
            vp = np.asarray(cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse), dtype=np.float32)
            view = make_view(camera, 1080, occlusion=2, triangle_budget=8_000_000, threshold_range=(0.5, 3.0))
            torch.cuda.synchronize()
            t = [time.perf_counter()]
            if f:
                scene.hzb_from_cut(0, vp, w, h, with_phase2=True)
                scene.sync()
            t.append(time.perf_counter())
            scene.cut([view])
            scene.sync()
            t.append(time.perf_counter())
            scene.hzb_from_cut(0, vp, w, h)
            scene.sync()
            t.append(time.perf_counter())
            scene.cut_phase2([view])
            scene.sync()
            t.append(time.perf_counter())
            counters, _, _, met = scene.read(0, max_entries=1)
            c2, _ = scene.read_phase2(0)
            if lap:
                frames.append({"ms": (t[4] - t[0]) * 1e3, "raster_ms": (t[1] - t[0] + t[3] - t[2]) * 1e3,
                               "cut_ms": (t[2] - t[1] + t[4] - t[3]) * 1e3, "tests": int(met.clusters_tested[0] + met.clusters_tested[1]),
                               "drawn": int(c2[2]), "added": int(c2[2]) - int(counters[2])})
    cut_s = sum(fr["cut_ms"] for fr in frames) / 1e3
    out["config2_world_flythrough"] = {
        "frames": len(frames), "frame_ms_p50": percentile([fr["ms"] for fr in frames], 50),
        "frame_ms_p99": percentile([fr["ms"] for fr in frames], 99),
        "cut_ms_p50": percentile([fr["cut_ms"] for fr in frames], 50), "cut_ms_p99": percentile([fr["cut_ms"] for fr in frames], 99),
        "raster_ms_mean": float(np.mean([fr["raster_ms"] for fr in frames])),
        "views_clusters_per_s": sum(fr["tests"] for fr in frames) / cut_s, "visible_clusters_per_s": sum(fr["drawn"] for fr in frames) / cut_s,
        "cluster_tests_per_frame": float(np.mean([fr["tests"] for fr in frames])), "drawn_per_frame": float(np.mean([fr["drawn"] for fr in frames])),
        "phase2_added_per_frame": float(np.mean([fr["added"] for fr in frames]))}
    log(f"config 2 ({time.time() - t0:.1f} s): {out['config2_world_flythrough']}")
    scene.close()
    return out


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported: the cut of one scene runs on one GPU"}))
        sys.exit(2)
    import torch

    from vgcut import camera as cam
    from vgcut import native, scenegen, scenes
    from vgcut.virtual_geometry import make_view

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.zeros(1, device="cuda:0")
    if args.plain_configs:
        if args.profile_only:
            # A plain 1 GiB device copy first: its FETCH_SIZE / WRITE_SIZE calibrate the counters' units.
            src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
        print(json.dumps({"plain_configs": plain_configs_bench(args.cpu_seconds)}))
        return
    t0 = time.time()
    spec = scenes.world(mesh_count=args.meshes, instances=args.instances, terrain_grid=512)
    log(f"world: {spec.instance_count} instances, {len(spec.meshes)} meshes, "
        f"{sum(d.meshletCount for d in spec.meshes)} meshlets, {spec.full_detail_triangles / 1e9:.1f}B full-detail "
        f"triangles ({time.time() - t0:.1f} s)")

    if args.serve_only:
        print(json.dumps({"serving": serve_bench(spec.meshes[:spec.props], make_view, cam, scenegen, SERVE_RATES, args.serve_seconds, args.max_meshlets)}))
        return
    if args.expand_only:
        vd = scenes.mixed_views(spec, args.views)
        schedule = tuple(int(x) for x in args.schedule.split(",")) if args.schedule else (12, 10, 5)
        print(json.dumps({"expand": expand_bench(spec, vd, args.steps, args.warmup, schedule, args.profile_only)}))
        return

    views_n = args.views
    ctx = native.Context(max_triangles=MAX_TRIANGLES, max_shadow_triangles=MAX_SHADOW_TRIANGLES, max_meshlets=args.max_meshlets,
                         max_views=views_n, mesh_slots=72, device=0)
    scene = native.Scene(ctx)
    t0 = time.time()
    scenes.load_into(scene, spec, occlusion=True)
    vd = scenes.mixed_views(spec, views_n)
    views = [v for v, _ in vd]
    for v, depth in vd:
        if depth is not None:
            scene.set_hzb(v.slot, depth)
    log(f"scene loaded, {sum(d is not None for _, d in vd)} HZBs ({time.time() - t0:.1f} s)")
    if args.two_phase_only:
        # A plain 1 GiB device copy first: its FETCH_SIZE / WRITE_SIZE calibrate the counters' units.
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print(json.dumps({"two_phase_occlusion": two_phase_bench(scene, vd, args.steps), "copy_bytes": 2 * (1 << 28) * 4}))
        return

    initial = np.zeros(8, dtype=np.uint32)
    initial[0] = initial[1] = 10000

    def reset_states():
        for v in views:
            scene.set_state(v.slot, initial)

    # Launch schedules: tuned on the batched views and, for the plain path, on the heaviest single view.
    if args.schedule:
        fixed = tuple(int(x) for x in args.schedule.split(","))
        scene.set_schedule(*fixed)
        batched_schedule = plain_schedule = fixed
    else:
        t0 = time.time()
        plain_schedule = scene.tune([views[0]], repeats=9)[0].as_tuple()
        batched_schedule = scene.tune(views, repeats=9)[0].as_tuple()
        log(f"tuned schedules: batched {batched_schedule}, plain {plain_schedule} ({time.time() - t0:.1f} s)")

    def step_work():
        """Cluster tests and drawn clusters of the last launch, and each view's metrics."""
        tested = selected = 0
        per_view = []
        for v in views:
            counters, _, _, met = scene.read(v.slot, max_entries=1)
            k = int(met.clusters_tested[0] + met.clusters_tested[1])
            s = int(counters[2]) + int(counters[7])
            tested += k
            selected += s
            per_view.append((k, s, met))
        return tested, selected, per_view

    # Warm-up (uploads, graph capture), then frames until the work repeats: each view's capacity controller
    # (GeometryPool.ts:602-640) moves its LOD scale frame over frame from the state at rest, so the first
    # frames test more clusters than the steady state (885M then 735M on the world). Every figure below is
    # taken in that steady state: the timed frames are the ones counted.
    for _ in range(max(1, args.warmup)):
        scene.cut(views, graph=True)
    scene.sync()
    tested = -1
    for _ in range(32):
        scene.cut(views, graph=True)
        scene.sync()
        previous = tested
        tested, selected, per_view = step_work()
        if tested == previous:
            break
    log(f"one step (steady state): {tested} cluster tests, {selected} clusters drawn over {views_n} views")
    steady_states = {v.slot: scene.state(v.slot).copy() for v in views}

    if args.profile_only:
        # A plain 1 GiB device copy first: its FETCH_SIZE / WRITE_SIZE calibrate the counters' units.
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        for _ in range(args.steps):
            scene.cut(views, graph=False)
        scene.sync()
        met = [scene.read(v.slot, max_entries=1)[3] for v in views]
        print(json.dumps({"profile_only": True, "steps": args.steps, "views": views_n, "copy_bytes": 2 * src.numel() * 4,
                          "schedule": batched_schedule,
                          "instances_tested": sum(int(m.instances_tested) for m in met),
                          "cluster_tests": tested, "clusters_drawn": selected}))
        return

    # ---------------- batched: one launch sequence for all views (graph replay) ----------------
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        scene.cut(views, graph=True)
    scene.sync()
    torch.cuda.synchronize()
    batched_s = (time.perf_counter() - t0) / args.steps
    tested_after = step_work()[0]
    if tested_after != tested:  # the steady state moved during the timed frames: count the smaller work
        log(f"work changed during the timed frames: {tested} then {tested_after} cluster tests")
        tested = min(tested, tested_after)

    # ---------------- plain: one view per launch sequence ----------------
    plain_s = None
    if not args.no_plain:
        scene.set_schedule(*plain_schedule)
        for v in views[:4]:
            scene.cut([v])
        scene.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for v in views:
                scene.cut([v])
        scene.sync()
        torch.cuda.synchronize()
        plain_s = (time.perf_counter() - t0) / args.steps

    # ---------------- latency: launch to completion ----------------
    lat_batched, lat_plain = [], []
    scene.set_schedule(*batched_schedule)
    for i in range(args.latency_samples):
        t = time.perf_counter()
        scene.cut(views, graph=True)
        scene.sync()
        lat_batched.append((time.perf_counter() - t) * 1e3)
    if not args.no_plain:
        scene.set_schedule(*plain_schedule)
        for i in range(args.latency_samples):
            v = views[i % views_n]
            t = time.perf_counter()
            scene.cut([v])
            scene.sync()
            lat_plain.append((time.perf_counter() - t) * 1e3)
    scene.set_schedule(*batched_schedule)

    # ---------------- per-kernel times (HIP events on the library's stream) ----------------
    scene.profile(True)
    prof_steps = max(3, args.steps // 2)
    for _ in range(prof_steps):
        scene.cut(views, graph=False)
    scene.sync()
    kernels, seqs = scene.profile_read()
    scene.profile(False)
    kernel_ms = {name: ms / max(1, seqs) for name, ms in kernels}
    dominant = max(kernel_ms, key=kernel_ms.get)

    # ---------------- algorithmic bytes (unique inputs per launch, outputs per view; steady state) ----------------
    inst_unique, meshlet_unique = scene.footprint(views)
    cells = len(scene.pool(native.VG_POOL_CELLS)) // 8
    levels = sum(d.levelCount for d in spec.meshes)
    W = sum(int(m.work_items[0] + m.work_items[1]) for _, _, m in per_view)
    O = sum(int(m.occlusion_tests) for _, _, m in per_view)
    slots = scene.mesh_slots
    hzb_bytes = sum(4 * d.shape[0] * d.shape[1] * (2 + 4 / 3) for _, d in vd if d is not None)
    per_kernel_bytes = {
        "k_cells": 32 * cells * views_n,
        "k_instances": 64 * inst_unique + 32 * levels + 16 * W,
        "k_meshlets": 64 * meshlet_unique + 16 * W + 32 * selected + 16 * O,
        "k_prefix(check 1)": 40 * slots * views_n,
    }
    launch_bytes = 32 * cells + 64 * inst_unique + 64 * meshlet_unique + 32 * levels + 32 * W + 32 * selected + 16 * O + 40 * slots * views_n
    dom_bytes = per_kernel_bytes.get(dominant, launch_bytes)
    dom_ms = kernel_ms[dominant]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic, traffic_source = measured_traffic(dominant)

    # ---------------- CPU baseline: the reference's CPU cut restated in Node, and the C oracle beside it ----------------
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle

        pool = oracle.pool_from_scene(scene, ctx.config.max_meshlets, MAX_TRIANGLES, MAX_SHADOW_TRIANGLES)
        scales = [steady_states[v.slot][0] / 10000.0 for v in views]
        threads = max(1, min(16, os.cpu_count() or 1))  # the box's CPU share
        t0 = time.time()
        node = oracle.node_cpu_baseline(pool, views, scales, seconds=args.cpu_seconds, threads=threads)
        one, many = node["single_thread"], node["threads"]
        log(f"CPU baseline (Node {node['node']}, {node['cpu_model']}, {node['logical_cpus']} logical CPUs): one thread "
            f"{one['tests_per_s'] / 1e6:.2f}M cluster tests/s over {one['views']} views, {many['n']} threads "
            f"{many['tests_per_s'] / 1e6:.1f}M/s ({time.time() - t0:.1f} s)")
        # The C oracle (binary32, the GPU's expression order, occlusion included), one thread, about 6 s of views.
        hzbs = {v.slot: oracle.hzb_build(d, 1024) for v, d in vd if d is not None}
        done_tests = used = 0
        t0 = time.perf_counter()
        for v in views:
            r = oracle.cut(pool, v, state=steady_states[v.slot], hzb=hzbs.get(v.slot))
            done_tests += int(r.stats.clusters_tested[0] + r.stats.clusters_tested[1])
            used += 1
            if time.perf_counter() - t0 > args.cpu_seconds / 2:
                break
        c_s = time.perf_counter() - t0
        cpu = {"value": many["tests_per_s"], "unit": "views*clusters tested/s", "cores": many["n"], "kind": "port",
               "cpu_model": node["cpu_model"], "logical_cpus": node["logical_cpus"], "runtime": f"node {node['node']}",
               "sample": (f"oracle/cpu_cut.js: the reference's CPU cut (cut.ts, fp64, Math.hypot, view*model) with the GPU pass's cell "
                          f"and instance culls, camera cut without occlusion, at each view's steady-state LOD scale; the first "
                          f"{many['views']} of the {views_n} views, {many['n']} worker threads sharded by cells, median of 3 after a "
                          f"warm-up: {many['tests']} cluster tests in {many['seconds']:.2f} s"),
               "single_thread": {"value": one["tests_per_s"], "views": one["views"], "cluster_tests": one["tests"],
                                 "seconds": one["seconds"]},
               "oracle_c": {"value": done_tests / c_s, "cores": 1, "views": used, "cluster_tests": done_tests, "seconds": c_s,
                            "what": "oracle/vgoracle.c: binary32 in the GPU kernels' order, all culling, occlusion, retries"}}
        del pool

    if args.check:
        from oracle import oracle

        pool = oracle.pool_from_scene(scene, ctx.config.max_meshlets, MAX_TRIANGLES, MAX_SHADOW_TRIANGLES)
        reset_states()
        scene.cut(views, graph=True)
        scene.sync()
        bad = overflowing = 0
        for v, d in vd:
            ref = oracle.cut(pool, v, hzb=None if d is None else oracle.hzb_build(d, 1024))
            counters, args_, draws, _ = scene.read(v.slot)
            over = max(int(ref.counters[2]), int(ref.counters[7])) > ctx.config.max_meshlets
            overflowing += over
            same = np.array_equal(counters, ref.counters) and np.array_equal(args_, ref.args)
            if same and not over:
                for cut in (0, 1):
                    n = int(ref.counters[2 + cut * 5])
                    g = draws[cut, :n][:, [0, 1, 3]]
                    r = ref.draws[cut, :n][:, [0, 1, 3]]
                    same = same and np.array_equal(g[np.lexsort(g.T[::-1])], r[np.lexsort(r.T[::-1])])
            if not same and not over:
                bad += 1
                log(f"check: view slot {v.slot} differs (selected gpu {counters[2]} oracle {ref.counters[2]})")
        log(f"check: {bad} of {views_n} views differ from the oracle; {overflowing} overflow the draw list "
            f"(their draw order, hence firstIndex and per-mesh counts, is not determined)")

    two_phase = two_phase_bench(scene, vd, args.steps)
    expand = None if args.no_expand else expand_bench(spec, vd, args.steps, args.warmup, batched_schedule)

    value = tested / batched_s
    out = {
        "metric": "views*clusters tested/s (batched, 64 mixed views, 3M-instance world)",
        "value": value,
        "unit": "views*clusters tested/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": batched_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"world {spec.instance_count} instances, {len(spec.meshes)} meshes, {views_n} mixed views",
                   "instances": spec.instance_count, "views": views_n, "error_threshold_px": 1.0},
        "plain": None if plain_s is None else {
            "views_clusters_per_s": tested / plain_s, "ms_per_step": plain_s * 1e3, "launches_per_step": views_n,
            "batched_speedup": plain_s / batched_s,
            "latency_ms_p50": percentile(lat_plain, 50), "latency_ms_p99": percentile(lat_plain, 99),
            "schedule": dict(zip(("instance_blocks_per_cu", "meshlet_blocks_per_cu", "meshlet_waves_per_simd"), plain_schedule))},
        "batched": {"latency_ms_p50": percentile(lat_batched, 50), "latency_ms_p99": percentile(lat_batched, 99),
                    "visible_clusters_per_s": selected / batched_s, "views_per_s": views_n / batched_s,
                    "schedule": dict(zip(("instance_blocks_per_cu", "meshlet_blocks_per_cu", "meshlet_waves_per_simd"),
                                         batched_schedule))},
        "work_per_step": {"cluster_tests": tested, "clusters_drawn": selected, "work_items": W, "occlusion_tests": O,
                          "unique_instances": inst_unique, "unique_meshlets": meshlet_unique, "cells": cells},
        # Each view's camera LOD scale in the capacity controller's steady state (its effective threshold is
        # error_threshold_px times this), in slot order.
        "steady_state_lod_scale": [round(int(steady_states[v.slot][0]) / 10000.0, 4) for v in views],
        "kernel_ms": kernel_ms,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
                     "algorithmic_bytes": dom_bytes,
                     "launch_algorithmic_bytes": launch_bytes, "kernel_ms": dom_ms},
        "cpu_baseline": cpu,
        "two_phase_occlusion": two_phase,
        "expand": expand,
        "serving": None if args.no_serve or args.profile_only else serve_bench(spec.meshes[:spec.props], make_view, cam, scenegen, SERVE_RATES,
                                                                               args.serve_seconds, args.max_meshlets),
    }
    print(json.dumps(out))
    # The metric in one line at the end of the log (the JSON line above is long).
    srv = out["serving"]
    serve_txt = ""
    if srv:
        rows = srv["rates"]
        serve_txt = (f"; served capacity {srv['batched_capacity_req_per_s']:.0f} req/s (plain {srv['plain_capacity_req_per_s']:.0f}), "
                     f"light-class p99 < 2 ms up to {srv['batched_light_p99_under_2ms_max_offered_req_per_s']} req/s offered; " + ", ".join(
            f"{r['offered_req_per_s']} req/s p50/p99 {r['batched']['latency_ms_p50']:.2f}/{r['batched']['latency_ms_p99']:.2f} ms"
            f" light p99 {r['batched']['light_p99_ms']:.2f} (fifo {r['batched_fifo']['latency_ms_p99']:.2f}, plain "
            f"{r['plain']['latency_ms_p50']:.2f}/{r['plain']['latency_ms_p99']:.2f})" for r in rows[:3]))
    plain = out["plain"]
    log(f"summary: batched {out['ms_per_step']:.3f} ms per {views_n} views ({value / 1e9:.1f}G views*clusters/s), latency p50/p99 "
        f"{out['batched']['latency_ms_p50']:.3f}/{out['batched']['latency_ms_p99']:.3f} ms"
        + ("" if plain is None else f"; plain {plain['ms_per_step']:.3f} ms ({plain['views_clusters_per_s'] / 1e9:.1f}G/s), "
                                    f"one-view latency p50/p99 {plain['latency_ms_p50']:.3f}/{plain['latency_ms_p99']:.3f} ms, "
                                    f"batched speedup {plain['batched_speedup']:.2f}x")
        + f"; {dominant} {achieved:.0f} GB/s = {achieved / HBM_PEAK_GBS:.3f} of HBM peak"
        + ("" if cpu is None else f"; CPU baseline {cpu['value'] / 1e6:.1f}M/s on {cpu['cores']} threads")
        + serve_txt)


if __name__ == "__main__":
    main()
