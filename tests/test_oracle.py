"""Pins the fp32 oracle (vgoracle.c) without retyping it.

- fp64 form of the reference's CPU cut (cut.ts:26-77): on one identity instance with culling off, the
  oracle's drawn clusters and blend ranges equal selectCut / selectBlendCut except clusters within
  1e-6 relative of a decision boundary; on many instances the built-in fp64 check finds no other
  mismatch;
- closed forms: known-answer projected errors placed 1e-3 relative off the threshold, sphere/plane
  frustum cases, the HZB pyramid against a brute-force max, occlusion conservativeness, log against
  libm, the capacity controller's documented behaviour;
- invariants: cells only cull whole cells whose instances would all be culled (inflated cells give the
  same cut), every drawn cut is complete (restated coverage check);
- planted mistakes fail: a transposed view matrix, a dropped draw entry, a removed cluster.
"""
import math

import numpy as np
import pytest

from helpers import draw_set
from oracle import cut64, oracle
from vgcut import camera as cam
from vgcut import native, scenegen, scenes
from vgcut.virtual_geometry import make_view


def host_pool(spec, max_meshlets=1 << 16, max_triangles=2_000_000, max_shadow=1_000_000, occlusion=False):
    ctx = native.Context(max_triangles=max_triangles, max_shadow_triangles=max_shadow, max_meshlets=max_meshlets, device=-1)
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=occlusion)
    pool = oracle.pool_from_scene(scene, max_meshlets, max_triangles, max_shadow)
    return ctx, scene, pool


@pytest.fixture(scope="module")
def terrain_pool():
    d = scenegen.reference_terrain()
    spec = scenes.SceneSpec([d], [np.eye(4, dtype=np.float32).reshape(1, 16)], [1e30], [0.0], [False])
    ctx, scene, pool = host_pool(spec)
    yield d, pool
    scene.close()
    ctx.close()


def _cpu_view(v, model=None):
    vm = np.array(v.view_matrix[:], dtype=np.float64)
    return {"viewMatrix": vm, "projScale": float(v.proj_scale), "near": float(v.near_plane)}


@pytest.mark.parametrize("threshold", [0.5, 1.0, 3.0])
def test_single_instance_equals_select_cut(terrain_pool, threshold):
    """Special case with a textbook answer: identity instance, no culling = selectCut (cut.ts:26-48)."""
    d, pool = terrain_pool
    rand = cut64.park_miller(7)
    for _ in range(8):
        eye = ((rand() - 0.5) * 300, 2 + rand() * 80, (rand() - 0.5) * 300)
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 5000).look_at(eye, ((rand() - 0.5) * 200, 0, (rand() - 0.5) * 200))
        v = make_view(camera, 1080, error_threshold=threshold, frustum_culling=False, min_pixel_radius=0.0, lod_blend=1.0,
                      max_error_threshold=1e6)
        r = oracle.cut(pool, v, fp64=True)
        assert r.stats.mismatch_hard == 0
        n = int(r.counters[2])
        got = np.zeros(d.meshletCount, dtype=np.uint8)
        got[r.draws[0, :n, 1]] = 1
        want = cut64.select_cut(d, _cpu_view(v), threshold)
        diff = np.nonzero(got != want)[0]
        assert len(diff) <= r.stats.near_threshold, f"{len(diff)} differences, {r.stats.near_threshold} near-threshold"
        assert not cut64.verify_cut_coverage(d, got) and not cut64.verify_cut_overlap(d, got)


def test_blend_ranges_equal_select_blend_cut(terrain_pool):
    """At 3 px the band is lodBlend wide (smoothstep saturates): blend ranges equal selectBlendCut's."""
    d, pool = terrain_pool
    for eye in ((0, 40, 260), (150, 6, -40), (-300, 200, 300)):
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 5000).look_at(eye, (0, 0, 0))
        v = make_view(camera, 1080, error_threshold=3.0, frustum_culling=False, min_pixel_radius=0.0, lod_blend=2.0,
                      max_error_threshold=1e6)
        r = oracle.cut(pool, v, fp64=True)
        assert r.stats.mismatch_hard == 0
        lo = np.zeros(d.meshletCount, dtype=np.int64)
        hi = np.zeros(d.meshletCount, dtype=np.int64)
        n = int(r.counters[2])
        e = r.draws[0, :n]
        lo[e[:, 1]] = (e[:, 3] >> 8) & 255
        hi[e[:, 1]] = (e[:, 3] >> 16) & 255
        wlo, whi = cut64.select_blend_cut(d, _cpu_view(v), 3.0, 2.0)
        bad = np.nonzero((lo != wlo) | (hi != whi))[0]
        assert len(bad) <= r.stats.near_threshold
        assert len(bad) <= 2


def test_many_instances_match_fp64(gpu_free_ruins):
    ctx, scene, pool = gpu_free_ruins
    rng = np.random.default_rng(9)
    checked = 0
    for k in range(6):
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at(
            (rng.uniform(-50, 50), rng.uniform(1.5, 40), rng.uniform(-50, 50)), (rng.uniform(-60, 60), 0, rng.uniform(-60, 60)))
        for t in (0.5, 1.0, 2.0, 3.0):
            r = oracle.cut(pool, make_view(camera, 1080, error_threshold=t, shadow_sweep=(10.0, -30.0, 5.0)), fp64=True)
            assert r.stats.mismatch_hard == 0
            # The flag covers the fade-step boundaries too: each drawn meshlet's two step values fall within
            # 1e-6 * fadeScale of a half-step with probability 2 * 2e-6 * 255 / ln(2) ~ 1.5e-3, and siblings
            # that share their LOD inputs are flagged together.
            assert r.stats.near_threshold <= 4e-3 * max(1, r.stats.lod_checked)
            checked += r.stats.lod_checked
    assert checked > 100_000


@pytest.fixture(scope="module")
def gpu_free_ruins():
    spec = scenes.ruins(mesh_count=4, tri_range=(8_000, 40_000), instances=1500, half_extent=60.0, seed=17)
    spec.cast_shadow = [True, False, True, False]
    ctx, scene, pool = host_pool(spec)
    yield ctx, scene, pool
    scene.close()
    ctx.close()


def test_inflated_cells_give_the_same_cut(gpu_free_ruins):
    """Cell culling is conservative: with every cell sphere 10x larger the cut is identical."""
    ctx, scene, pool = gpu_free_ruins
    fat = oracle.Pool(**{**pool.__dict__, "cells": pool.cells.copy()})
    c = fat.cells.reshape(-1, 8)
    c[:, 3] *= 10
    rng = np.random.default_rng(4)
    for _ in range(5):
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at(
            (rng.uniform(-50, 50), rng.uniform(1.5, 20), rng.uniform(-50, 50)), (rng.uniform(-60, 60), 0, rng.uniform(-60, 60)))
        v = make_view(camera, 1080, min_pixel_radius=2.0)
        a, b = oracle.cut(pool, v), oracle.cut(fat, v)
        assert np.array_equal(a.counters[2:], b.counters[2:])
        for cut in (0, 1):
            n = int(a.counters[2 + cut * 5])
            assert np.array_equal(draw_set(a.draws[cut], n), draw_set(b.draws[cut], n))
        assert a.counters[0] <= b.counters[0]


def _one_meshlet_dag(error: float, parent_error: float, center=(0.0, 0.0, 0.0), radius=1.0):
    """Two clusters: a leaf (error e, parent error p) and its root (error p); spheres at `center`."""
    b = np.zeros((2, 16), dtype=np.float32)
    b[0, 0:4] = (*center, radius)
    b[0, 4:8] = (*center, radius)
    b[0, 8:10] = (error, parent_error)
    b[0, 12:16] = (*center, radius)
    b[1, 0:4] = (*center, radius)
    b[1, 8:10] = (parent_error, 1e30)
    b[1, 12:16] = (*center, radius)
    info = np.array([[128, 0, 0, 0], [64, 384, 1, 0]], dtype=np.uint32)
    return scenegen.VirtualMeshData(
        meshletCount=2, meshletBounds=b.reshape(-1), meshletInfo=info.reshape(-1),
        levelRanges=np.array([0, 1, 1, 1], dtype=np.uint32),
        levelErrors=np.array([error, parent_error, parent_error, 1e30], dtype=np.float32),
        levelGuardErrors=np.array([error, parent_error, parent_error, 1e30], dtype=np.float32),
        replacementGroup=np.array([0, -1], dtype=np.int32), replacementStart=np.array([0, 1], dtype=np.uint32),
        replacementIndices=np.array([1], dtype=np.uint32),
        boundingSphere=np.array([*center, radius], dtype=np.float32), lodBoundsSphere=np.array([*center, radius * 1.0001], dtype=np.float32),
        cullBoundsSphere=np.array([*center, radius * 1.0001], dtype=np.float32), leafTriangles=128, leafMeshlets=1,
        rootTriangles=64, rootMeshlets=1)


@pytest.mark.parametrize("rel", [-1e-3, 1e-3])
def test_known_answer_projected_error(rel):
    """Closed form: px = e * projScale / (D - r). A leaf (error 0) and its root (error p) share a sphere of
    radius r; a camera at D where the root projects to t (1 + rel) px draws the root (rel < 0) or the
    leaf (rel > 0)."""
    p, r, t = 0.05, 1.0, 1.0
    camera0 = cam.perspective_camera(60, 16 / 9, 0.01, 1e5)
    ps = camera0.projectionMatrix[5] * 1080 * 0.5
    D = p * ps / (t * (1 + rel)) + r
    d = _one_meshlet_dag(0.0, p, radius=r)
    spec = scenes.SceneSpec([d], [np.eye(4, dtype=np.float32).reshape(1, 16)], [1e30], [0.0], [False])
    ctx, scene, pool = host_pool(spec)
    camera = cam.perspective_camera(60, 16 / 9, 0.01, 1e5).look_at((0, 0, D), (0, 0, 0))
    v = make_view(camera, 1080, error_threshold=t, min_pixel_radius=0.0)
    res = oracle.cut(pool, v, fp64=True)
    drawn = set(res.draws[0, :int(res.counters[2]), 1].tolist())
    assert drawn == ({1} if rel < 0 else {0})
    assert res.stats.near_threshold == 0 and res.stats.mismatch_hard == 0


def test_frustum_closed_form():
    """A unit sphere just inside / just outside the left plane of a camera is kept / culled."""
    d = _one_meshlet_dag(0.0, 1e30)
    camera = cam.perspective_camera(90, 1.0, 0.1, 1000).look_at((0, 0, 0), (0, 0, -1))
    # 90 degree frustum: the left plane is x = z (x = -d at depth d). A sphere of radius 1 at depth 20 touches
    # it when its centre is sqrt(2) left of the edge.
    for dx, kept in ((-20 - math.sqrt(2) + 0.01, True), (-20 - math.sqrt(2) - 0.01, False)):
        m = np.eye(4, dtype=np.float32)
        m[3, 0], m[3, 2] = dx, -20.0
        spec = scenes.SceneSpec([d], [m.reshape(1, 16)], [1e30], [0.0], [False])
        ctx, scene, pool = host_pool(spec)
        res = oracle.cut(pool, make_view(camera, 1080, min_pixel_radius=0.0))
        assert (int(res.counters[2]) > 0) == kept


def test_hzb_pyramid_is_the_brute_force_max():
    rng = np.random.default_rng(3)
    for w, h in ((960, 540), (37, 1), (1, 1), (64, 64), (1023, 77)):
        depth = rng.uniform(0, 1, (h, w)).astype(np.float32)
        hzb = oracle.hzb_build(depth, 1024)
        levels = int(hzb[64])
        assert hzb[1] == w and hzb[2] == h
        for lv in range(levels):
            off, lw, lh = int(hzb[lv * 4]), int(hzb[lv * 4 + 1]), int(hzb[lv * 4 + 2])
            mip = hzb[off:off + lw * lh].reshape(lh, lw)
            f = 1 << lv
            for y in range(0, lh, max(1, lh // 7)):
                for x in range(0, lw, max(1, lw // 7)):
                    want = depth[y * f:min(h, (y + 1) * f), x * f:min(w, (x + 1) * f)].max()
                    assert mip[y, x] == want, (w, h, lv, x, y)
        assert int(hzb[(levels - 1) * 4 + 1]) == 1 and int(hzb[(levels - 1) * 4 + 2]) == 1


def test_occlusion_is_conservative():
    """Whenever occluded() is true, every mip-0 texel covered by the sphere's projection is nearer than it."""
    rng = np.random.default_rng(8)
    camera = cam.perspective_camera(60, 16 / 9, 0.1, 1000).look_at((0, 0, 0), (0, 0, -1))
    P = camera.projectionMatrix.astype(np.float32)
    w, h = 320, 180
    depth = np.full((h, w), 1.0, dtype=np.float32)
    depth[:, :200] = 0.99  # an occluder wall on the left
    hzb = oracle.hzb_build(depth, 1024)
    hidden = 0
    for _ in range(3000):
        c = np.array([rng.uniform(-30, 30), rng.uniform(-20, 20), -rng.uniform(1, 200)], dtype=np.float32)
        r = float(rng.uniform(0.1, 5))
        if not oracle.occluded(P, 0.1, hzb, c, r):
            continue
        hidden += 1
        # Closed-form screen bounds of the sphere's box and its nearest depth.
        nd, fd = -c[2] - r, -c[2] + r
        xs = [(c[0] + sx * r) / dd * P[0] - P[8] for sx in (-1, 1) for dd in (nd, fd)]
        ys = [(c[1] + sy * r) / dd * P[5] - P[9] for sy in (-1, 1) for dd in (nd, fd)]
        # Texel footprint clamped to the screen (parts off screen are not visible, OcclusionCulling.ts:147-150).
        clamp = lambda v, hi: min(hi, max(0, v))  # noqa: E731
        x0 = clamp(int(math.floor((min(xs) * 0.5 + 0.5) * w)), w - 1)
        x1 = clamp(int(math.floor((max(xs) * 0.5 + 0.5) * w)), w - 1)
        y0 = clamp(int(math.floor((0.5 - max(ys) * 0.5) * h)), h - 1)
        y1 = clamp(int(math.floor((0.5 - min(ys) * 0.5) * h)), h - 1)
        min_z = P[14] / nd - P[10]
        assert depth[y0:y1 + 1, x0:x1 + 1].max() < min_z
    assert hidden > 50


def test_log_matches_libm():
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), [1.0, 2.0, 0.5, 1e-30, 1e30, 1.0009765625, 0.99999994]]).astype(np.float32)
    for x in xs:
        got = oracle.logf(float(x))
        want = math.log(float(x))
        ulp = np.spacing(np.float32(abs(want)) if want != 0 else np.float32(1e-45))
        assert abs(got - want) <= 1.0 * float(ulp) + 1e-45, (x, got, want)
    assert oracle.logf(1.0) == 0.0


def test_capacity_controller_behaviour(gpu_free_ruins):
    """GeometryPool.ts:1023-1069: an overflowing cut is redone until it fits or the scale reaches the limit;
    with room the scale glides back by 1% per frame, never below 1."""
    ctx, scene, pool0 = gpu_free_ruins
    pool = oracle.Pool(**{**pool0.__dict__, "cut_capacity": (30_000, 30_000), "max_meshlets": 1 << 12})
    camera = cam.perspective_camera(60, 16 / 9, 0.1, 2000).look_at((0, 3, -40), (0, 0, 20))
    v = make_view(camera, 1080, error_threshold=0.5)
    r = oracle.cut(pool, v)
    assert r.stats.retries[0] >= 1 and r.state[0] > 10000
    # At most 2 retry rounds per frame (each at least 1.1x coarser); between frames the steer keeps
    # coarsening while the fill is above 0.9, so within a few frames the cut fits or hits the limit.
    for _ in range(12):
        scale = r.state[0] / 10000
        fill = max(r.counters[6] / 30_000, r.counters[2] / 4096)
        if fill <= 1.0 or scale >= v.lod_scale_limit * 0.999:
            break
        before = scale
        r = oracle.cut(pool, v, state=r.state)
        assert r.state[0] / 10000 >= before * 1.1 ** min(1, int(r.stats.retries[0])) * 0.9999
    assert fill <= 1.0 or scale >= v.lod_scale_limit * 0.999
    # With ample room the steer brings the scale down 1% per frame (GeometryPool.ts:1062-1066).
    big = oracle.Pool(**{**pool0.__dict__, "cut_capacity": (10**9, 10**9), "max_meshlets": 1 << 16})
    state = r.state.copy()
    for _ in range(3):
        before = state[0] / 10000
        state = oracle.cut(big, v, state=state).state
        assert abs(state[0] / 10000 - max(1.0, before / 1.01)) <= 1e-4


def test_planted_mistakes_are_caught(terrain_pool, gpu_free_ruins):
    """A transposed view matrix (a plausible index mistake) is caught by the fp64 check; a dropped draw
    entry is caught by the set comparison; a removed cluster is caught by the coverage check."""
    d, pool = terrain_pool
    camera = cam.perspective_camera(60, 16 / 9, 0.1, 50000).look_at((300, 400, 2000), (0, 0, 0))
    v = make_view(camera, 1080, frustum_culling=False, min_pixel_radius=0.0)
    vt = make_view(camera, 1080, frustum_culling=False, min_pixel_radius=0.0)
    m = np.array(v.view_matrix[:]).reshape(4, 4).T.reshape(16)
    for i in range(16):
        vt.view_matrix[i] = float(m[i])
    good = oracle.cut(pool, v, fp64=True)
    assert good.stats.mismatch_hard == 0
    # The fp64 form keeps the right matrix: run the mistaken fp32 cut and check its set against it.
    bad = oracle.cut(pool, vt)
    n = int(bad.counters[2])
    got = np.zeros(d.meshletCount, dtype=np.uint8)
    got[bad.draws[0, :n, 1]] = 1
    assert np.count_nonzero(got != cut64.select_cut(d, _cpu_view(v), 1.0)) > 0
    n = int(good.counters[2])
    a = draw_set(good.draws[0], n)
    assert not np.array_equal(a, draw_set(good.draws[0][1:], n - 1))
    sel = np.zeros(d.meshletCount, dtype=np.uint8)
    sel[good.draws[0, :n, 1]] = 1
    assert not cut64.verify_cut_coverage(d, sel)
    # verifyCutCoverage works per group (any output of a replacing group covers) and a group's outputs may
    # merge into different coarser groups, so one dropped meshlet can leave every leaf covered through a
    # sibling. Drop every selected meshlet on leaf 0's paths to the coarsest level: leaf 0 must be reported.
    group, start, lst = d.replacementGroup, d.replacementStart, d.replacementIndices
    reach, todo = set(), [0]
    while todo:
        i = todo.pop()
        if i in reach:
            continue
        reach.add(i)
        if group[i] >= 0:
            todo.extend(int(k) for k in lst[start[group[i]]:start[group[i] + 1]])
    dropped = [i for i in reach if sel[i]]
    assert dropped
    sel[dropped] = 0
    assert 0 in cut64.verify_cut_coverage(d, sel)
