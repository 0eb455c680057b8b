"""Two-phase occlusion in the oracle (no GPU): the second phase against an HZB that hides nothing recovers
exactly the cut without occlusion; against the cut's own HZB it adds nothing; its regions start after
the cut's requested indices. These pin the oracle's restatement independently of its code."""
import numpy as np
import pytest

from helpers import ground_camera, small_ruins
from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenegen, scenes
from vgcut.virtual_geometry import make_view

MAX_MESHLETS = 1 << 16


@pytest.fixture(scope="module")
def pool():
    spec = small_ruins(seed=71, instances=3000, shadows=True)
    # capacities the cuts never fill: no capacity retry changes a band between the compared cuts
    ctx = native.Context(max_triangles=50_000_000, max_shadow_triangles=25_000_000, max_meshlets=MAX_MESHLETS, device=-1)
    scene = native.Scene(ctx)
    scenes.load_into(scene, spec, occlusion=True)
    return oracle.pool_from_scene(scene, MAX_MESHLETS, 50_000_000, 25_000_000)


def hzb_for(camera, seed, w=480, h=270):
    vp = cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse)
    depth = scenegen.heightfield_depth(vp, w, h, (0.0, 0.0), 120.0, 128, 6.0, 1 / 25.0, seed)
    return oracle.hzb_build(depth, 1024)


CAMERAS = [((0.0, 2.0, -70.0), (0.0, 0.0, 60.0)), ((50.0, 2.0, 10.0), (-60.0, 2.0, 0.0))]


def entry_set(draws, lo, hi):
    d = draws[0, lo:hi][:, [0, 1, 3]]
    return d[np.lexsort(d.T[::-1])]


@pytest.mark.parametrize("k", [0, 1])
def test_second_phase_without_occluders_recovers_the_unoccluded_cut(pool, k):
    camera = ground_camera(*CAMERAS[k])
    view = make_view(camera, 540, occlusion=2, error_threshold=[1.0, 2.0][k], shadow_sweep=(10.0, -30.0, 5.0))
    far = oracle.hzb_build(np.ones((270, 480), dtype=np.float32), 1024)
    r = oracle.cut_two_phase(pool, view, hzb_for(camera, 77 + k), far)
    assert r.rejects[0] + r.rejects[1] > 0, "the first HZB must hide something"
    plain_view = make_view(camera, 540, occlusion=False, error_threshold=[1.0, 2.0][k], shadow_sweep=(10.0, -30.0, 5.0))
    plain = oracle.cut(pool, plain_view)
    assert plain.stats.retries[0] == 0 and r.first.stats.retries[0] == 0
    n1 = int(r.first.counters[2])
    total = int(r.counters2[2])
    assert total == int(plain.counters[2])
    both = np.concatenate([r.draws[0, :total][:, [0, 1, 3]]])
    want = plain.draws[0, :total][:, [0, 1, 3]]
    assert np.array_equal(both[np.lexsort(both.T[::-1])], want[np.lexsort(want.T[::-1])])
    # per (mesh slot, part) index counts of the two phases add up to the unoccluded cut's
    slots = pool.slots
    c1 = r.first.counters[12:12 + slots * 4].reshape(slots, 2, 2)[:, :, 0]
    c2 = r.counters2[12:12 + slots * 4].reshape(slots, 2, 2)[:, :, 0]
    cp = plain.counters[12:12 + slots * 4].reshape(slots, 2, 2)[:, :, 0]
    assert np.array_equal(c1.astype(np.int64) + c2, cp)
    # the shadow cut is the first phase's alone
    assert np.array_equal(r.first.counters[7:12], plain.counters[7:12])
    assert n1 < total
    # second-phase regions follow the cut's requested indices
    requested1 = int(r.first.counters[2 + 4])
    starts = r.counters2[12 + 1:12 + slots * 4:2]
    counts = r.counters2[12:12 + slots * 4:2]
    assert int(starts[counts > 0].min()) == requested1 if (counts > 0).any() else True
    assert int(r.counters2[2 + 4]) == requested1 + int(counts.sum())


def test_second_phase_against_the_same_hzb_adds_nothing(pool):
    camera = ground_camera(*CAMERAS[0])
    view = make_view(camera, 540, occlusion=2)
    hzb = hzb_for(camera, 77)
    r = oracle.cut_two_phase(pool, view, hzb, hzb)
    assert r.rejects[0] + r.rejects[1] > 0
    assert int(r.counters2[2]) == int(r.first.counters[2])
    assert not r.counters2[12:12 + pool.slots * 8:2].any()
    # the cut itself equals a one-phase cut with occlusion
    one = oracle.cut(pool, make_view(camera, 540, occlusion=True), hzb=hzb)
    assert np.array_equal(r.first.counters, one.counters) and np.array_equal(r.first.args, one.args)


def test_planted_mistake_fails(pool):
    """Re-testing against the cut's HZB instead of the new one (the mistake) adds nothing where the
    unoccluded cut's extra clusters must come back."""
    camera = ground_camera(*CAMERAS[1])
    view = make_view(camera, 540, occlusion=2)
    hzb = hzb_for(camera, 78)
    far = oracle.hzb_build(np.ones((270, 480), dtype=np.float32), 1024)
    right = oracle.cut_two_phase(pool, view, hzb, far)
    wrong = oracle.cut_two_phase(pool, view, hzb, hzb)
    assert int(right.counters2[2]) != int(wrong.counters2[2])
