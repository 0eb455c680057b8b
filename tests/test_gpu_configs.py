"""Parity at the benchmark configurations' full sizes (through the C ABI, against the oracle):

- the ruins scene (113,106 instances of 48 props of 0.2M-5M triangles): 8 seeded poses at 2 m eye
  height, each at 1080p and 2160p, occlusion against the rasterized 30 m heightfield, the reference's
  default triangle capacities (so the capacity retries run) with a draw list of 8M entries, all 16 views
  in one launch. No round may overflow the draw list: an overflowed list keeps the entries appended
  first, so the requested-triangle counts and the retry's scale depend on the append order (in the
  reference too);
- the 3M-instance world with its terrain as a mesh: every fourth frame of the 16-frame seeded fly-through
  (4 frames; the oracle's budget histogram takes about 20 s per frame) at 1080p with an 8M-triangle budget
  (histogram pick; this world's cut exceeds 8M triangles even at 3 px, so the pick is clamped to the last
  threshold of thresholdRange, as the reference clamps) and two-phase occlusion against the previous
  frame's visible set: the first phase's HZB is the occluder raster of the previous frame's cut (both
  phases) with this frame's camera, the second phase's the raster of the first phase's clusters, each
  HZB equal to the oracle's raster of the oracle's own cut; the controller state carried over.

Counters (controller state included), draw args and drawn sets are compared bit for bit, as everywhere.
"""
import math

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


def test_world_flythrough_budget_two_phase(gpu_available):
    spec = scenes.world(terrain_grid=512)
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
        b, _, chosen = oracle.budget_pick(pool, view, budget, thresholds, state=state, hzb=hzb1)
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
    assert hidden > 0, "the previous frames' visible sets must hide something"
