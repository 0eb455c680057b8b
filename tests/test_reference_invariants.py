"""The reference's own cut tests, restated (tests/cut.test.ts:5-117, tests/fastpath.test.ts:6-176).

They run on the reference tests' terrain (PlaneGeometry(400, 400, 96, 96), y = 30 sin(0.05 x) cos(0.04 z))
with our DAG builder in place of meshoptimizer's, and with the fp64 CPU cut (oracle/cut64.py), cameras
from the same Park-Miller sequence (seeds 7 and 11) and three.js camera conventions.
"""
import math

import numpy as np
import pytest

from oracle import cut64
from vgcut import camera as cam
from vgcut import scenegen


@pytest.fixture(scope="module")
def terrain():
    return scenegen.reference_terrain()


def _view(camera, height=1080, scale=None, model=None):
    vm = camera.matrixWorldInverse if model is None else cam.mat_mul(camera.matrixWorldInverse, model)
    v = {"viewMatrix": vm, "projScale": camera.projectionMatrix[5] * height * 0.5, "near": camera.near}
    if scale is not None:
        v["scale"] = scale
    return v


def test_dag_links(terrain):
    assert len(terrain.replacementGroup) == terrain.meshletCount  # cut.test.ts:22


def test_cut_covers_every_leaf_once_and_range_contains_it(terrain):
    """cut.test.ts:5-65: no holes at 0.5/1/3 px from 12 seeded cameras; the level range holds the cut."""
    rand = cut64.park_miller(7)
    worst = 0
    for _ in range(12):
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 5000)
        eye = ((rand() - 0.5) * 300, 2 + rand() * 80, (rand() - 0.5) * 300)
        target = ((rand() - 0.5) * 200, 0, (rand() - 0.5) * 200)
        camera.look_at(eye, target)
        for threshold in (0.5, 1, 3):
            view = _view(camera)
            selected = cut64.select_cut(terrain, view, threshold)
            holes = cut64.verify_cut_coverage(terrain, selected)
            worst = max(worst, len(holes))
            assert not holes, f"holes at {threshold}px from {eye}"
            assert not cut64.verify_cut_overlap(terrain, selected), f"overlaps at {threshold}px"
            start, end = cut64.instance_meshlet_range(terrain, view, threshold)
            idx = np.nonzero(selected)[0]
            assert np.all((idx >= start) & (idx < end)), "level range contains every selected meshlet"
    # Far away the range shrinks to a few coarse levels (cut.test.ts:56-64).
    camera = cam.perspective_camera(60, 16 / 9, 0.1, 50000).look_at((0, 2000, 8000), (0, 0, 0))
    view = {"viewMatrix": camera.matrixWorldInverse, "projScale": camera.projectionMatrix[5] * 540, "near": 0.1}
    start, end = cut64.instance_meshlet_range(terrain, view, 1)
    assert end - start < terrain.meshletCount / 4


def test_blend_band_steps_are_complete_cuts(terrain):
    """cut.test.ts:67-117: every step of the band is a complete cut without overlaps; blend 1 = plain cut."""
    blended = 0
    for eye in ((0, 40, 260), (150, 6, -40), (-300, 200, 300)):
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 5000).look_at(eye, (0, 0, 0))
        view = {"viewMatrix": camera.matrixWorldInverse, "projScale": camera.projectionMatrix[5] * 540, "near": 0.1}
        for threshold, blend in ((1, 2), (0.7, 1.5), (2, 3)):
            lo, hi = cut64.select_blend_cut(terrain, view, threshold, blend)
            for t in (threshold, threshold * blend * 0.999):
                plain = cut64.select_cut(terrain, view, t)
                assert np.all((plain == 0) | (hi > 0)), "plain cut inside the blend selection"
            for step in range(cut64.LOD_FADE_STEPS):
                selected = ((lo <= step) & (step < hi)).astype(np.uint8)
                assert not cut64.verify_cut_coverage(terrain, selected), f"hole at step {step}"
                assert not cut64.verify_cut_overlap(terrain, selected), f"overlap at step {step}"
            blended += int(np.sum((hi > 0) & ((lo > 0) | (hi < cut64.LOD_FADE_STEPS))))
    assert blended > 0
    camera = cam.perspective_camera(60, 16 / 9, 0.1, 5000).look_at((0, 40, 260), (0, 0, 0))
    view = {"viewMatrix": camera.matrixWorldInverse, "projScale": camera.projectionMatrix[5] * 540, "near": 0.1}
    plain = cut64.select_cut(terrain, view, 1)
    lo, hi = cut64.select_blend_cut(terrain, view, 1, 1)
    assert np.array_equal(plain == 1, hi > 0)
    assert np.all((hi == 0) | ((lo == 0) & (hi == cut64.LOD_FADE_STEPS)))


def _dist(a, b):
    return math.sqrt(sum((float(x) - float(y)) ** 2 for x, y in zip(a[:3], b[:3])))


def test_fast_path_invariants(terrain):
    """fastpath.test.ts:26-57."""
    B = terrain.meshletBounds.reshape(-1, 16).astype(np.float64)
    C = terrain.cullBoundsSphere.astype(np.float64)
    assert np.all(np.linalg.norm(B[:, 12:15] - C[:3], axis=1) + B[:, 15] <= C[3]), "cullBoundsSphere encloses every culling sphere"
    L = terrain.lodBoundsSphere.astype(np.float64)
    assert np.all(np.linalg.norm(B[:, 0:3] - L[:3], axis=1) + B[:, 3] <= L[3])
    nonroot = B[:, 9] < 1e29
    assert np.all(np.linalg.norm(B[nonroot, 4:7] - L[:3], axis=1) + B[nonroot, 7] <= L[3]), "lodBoundsSphere encloses parents"
    for lv in range(terrain.levelCount):
        s, n = terrain.levelRanges[lv * 2], terrain.levelRanges[lv * 2 + 1]
        own = terrain.meshletBounds.reshape(-1, 16)[s:s + n, 8]
        par = terrain.meshletBounds.reshape(-1, 16)[s:s + n, 9]
        assert terrain.levelGuardErrors[lv * 2] == own.max() and terrain.levelGuardErrors[lv * 2 + 1] == par.min()


def _translate_scale(t, s):
    m = np.eye(4) * s
    m[3, 3] = 1
    m[:3, 3] = t
    return m.T.reshape(16).copy()


def test_fast_path_is_conservative(terrain):
    """fastpath.test.ts:59-176: whenever the fast path fires it equals selectCut (also with blend 2), including
    cameras swept across the distance-band edge of every meshlet."""
    rand = cut64.park_miller(11)
    fast = fast_single = fast_blend = violations = 0
    for k in range(40):
        d = np.array([rand() - 0.5, 0.05 + rand() * 0.6, rand() - 0.5])
        d /= np.linalg.norm(d)
        dist = math.pow(10, 4.3 + rand() * 2) if k < 30 else 100 * math.pow(10, rand() * 2)
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 1e8).look_at(d * dist, ((rand() - 0.5) * 200, 0, (rand() - 0.5) * 200))
        for j in range(5):
            s = [1, 0.5, 2, 3][j % 4] * (0.75 + rand() * 0.5)
            model = _translate_scale(((rand() - 0.5) * 600, (rand() - 0.5) * 60, (rand() - 0.5) * 600), s)
            for threshold in (0.5, 1, 3):
                view = _view(camera, scale=s, model=model)
                rng = cut64.instance_guaranteed_range(terrain, view, threshold)
                if not rng:
                    continue
                fast += 1
                fast_single += (rng[1] - rng[0]) == 1
                sel = np.nonzero(cut64.select_cut(terrain, view, threshold))[0]
                if not (len(sel) == rng[1] - rng[0] and np.all((sel >= rng[0]) & (sel < rng[1]))):
                    violations += 1
                brng = cut64.instance_guaranteed_range(terrain, view, threshold, 2)
                if not brng:
                    continue
                fast_blend += 1
                lo, hi = cut64.select_blend_cut(terrain, view, threshold, 2)
                in_fast = (np.arange(terrain.meshletCount) >= brng[0]) & (np.arange(terrain.meshletCount) < brng[1])
                if not np.array_equal(in_fast, hi > 0) or np.any(in_fast & ((lo != 0) | (hi != cut64.LOD_FADE_STEPS))):
                    violations += 1
    # Edge sweep (fastpath.test.ts:112-171).
    B = terrain.meshletBounds.reshape(-1, 16).astype(np.float64)
    C = terrain.lodBoundsSphere[:3].astype(np.float64)
    probe = cam.perspective_camera(60, 16 / 9, 0.1, 1e8)
    ps = probe.projectionMatrix[5] * 1080 * 0.5
    for i in range(terrain.meshletCount):
        for mode in ("own", "parent"):
            o = 0 if mode == "own" else 4
            c, r, e = B[i, o:o + 3], B[i, o + 3], B[i, 8 if mode == "own" else 9]
            if not (0 < e < 1e29):
                continue
            off = float(np.linalg.norm(c - C))
            u = (c - C) / off if off > 0 else np.array([0.0, 1.0, 0.0])
            u = u * (1 if mode == "own" else -1)
            for threshold in (0.5, 1, 3):
                for s in (1, 2.5):
                    Dc = (e * s * ps) / threshold + s * (off + r) if mode == "own" else (e * s * ps) / threshold - s * (off - r)
                    if not Dc > 0:
                        continue
                    for kk in range(0, 33, 8):
                        f = 0.996 + kk * 0.00025
                        T = np.array([(rand() - 0.5) * 600, (rand() - 0.5) * 60, (rand() - 0.5) * 600])
                        model = _translate_scale(T, s)
                        center = T + C * s
                        camera = cam.perspective_camera(60, 16 / 9, 0.1, 1e8).look_at(center + u * Dc * f, center)
                        view = _view(camera, scale=s, model=model)
                        rng = cut64.instance_guaranteed_range(terrain, view, threshold)
                        if not rng:
                            continue
                        fast += 1
                        fast_single += (rng[1] - rng[0]) == 1
                        sel = np.nonzero(cut64.select_cut(terrain, view, threshold))[0]
                        if not (len(sel) == rng[1] - rng[0] and np.all((sel >= rng[0]) & (sel < rng[1]))):
                            violations += 1
    assert fast > 0 and fast_single > 0, "the fast path fires for far instances (not vacuous)"
    assert fast_blend > 0, "the fast path also fires with LOD blending"
    assert violations == 0
