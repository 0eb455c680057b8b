"""The CPU baseline (oracle/cpu_cut.js, the reference's CPU cut restated in Node with the GPU pass's culls)
counts the same work as the cut it is timed beside: on scenes without occlusion or shadow casters, its
cluster tests and drawn clusters per view equal the oracle's camera cut, which the GPU matches bit for bit.
A view at a higher LOD scale (the capacity controller's state) and a blended band are included.
"""
import shutil

import numpy as np
import pytest

from helpers import ground_camera, small_ruins
from oracle import oracle
from test_oracle import host_pool
from vgcut import camera as cam
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="node is missing")


def test_node_cpu_cut_counts_equal_the_oracle():
    spec = small_ruins(seed=91, instances=3000, meshes=4)
    ctx, scene, pool = host_pool(spec, max_meshlets=1 << 20, max_triangles=200_000_000, max_shadow=10_000_000)
    rng = np.random.default_rng(91)
    views, scales = [], []
    for k in range(6):
        camera = ground_camera((rng.uniform(-60, 60), rng.uniform(2, 30), rng.uniform(-60, 60)),
                               (rng.uniform(-60, 60), 0.0, rng.uniform(-60, 60)))
        views.append(make_view(camera, [720, 1080, 2160][k % 3], slot=0, error_threshold=[1.0, 0.6, 2.0][k % 3]))
        scales.append([1.0, 1.0, 1.7][k % 3])
    ortho = cam.orthographic_camera(-50, 50, 50, -50, 0.1, 400).look_at((10.0, 150.0, 0.0), (0.0, 0.0, 0.0))
    views.append(make_view(ortho, 2048, slot=0, min_pixel_radius=0.0))
    scales.append(1.0)
    got = oracle.node_cpu_baseline(pool, views, scales, seconds=60.0, threads=2)
    one = got["single_thread"]
    passes = one["views"] // len(views)  # a short view list is cut again until a second has passed
    assert passes >= 1 and one["views"] == passes * len(views), "the one-thread run must cover every view, whole passes here"
    tests = drawn = 0
    for v, s in zip(views, scales):
        state = oracle.initial_state()
        state[0] = int(round(s * 10000))
        r = oracle.cut(pool, v, state=state)
        assert int(r.stats.retries[0]) == 0
        tests += int(r.stats.clusters_tested[0])
        drawn += int(r.counters[2])
    assert drawn > 1000
    assert one["tests"] == passes * tests and one["drawn"] == passes * drawn, (one, tests, drawn)
    assert got["threads"]["tests"] == tests and got["threads"]["drawn"] == drawn
    assert got["threads"]["n"] == 2 and got["cpu_model"]
    scene.close()
    ctx.close()
