"""Two-phase occlusion on the GPU (vg_cut with occlusion = 2, then vg_cut_phase2 after a new HZB) against
the oracle's restatement: the cut, the second phase's counters and draw args bit-exact, each phase's
drawn set equal, the second phase's firstIndex ranges tiling its regions, and its expanded indices equal
to the oracle's expansion of the GPU's own draw list. Properties without the oracle: against an HZB that
hides nothing the two phases draw exactly the cut without occlusion; against the same HZB the second
phase adds nothing."""
import numpy as np
import pytest

from helpers import compare_cut, ground_camera, load_gpu, small_ruins, tiling_ok
from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenegen
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu

CAMERAS = [((0.0, 2.0, -70.0), (0.0, 0.0, 60.0)), ((50.0, 2.0, 10.0), (-60.0, 2.0, 0.0)), ((-40.0, 3.0, 40.0), (30.0, 0.0, -20.0))]
W, H = 480, 270


def depth_for(camera, seed):
    vp = cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse)
    return scenegen.heightfield_depth(vp, W, H, (0.0, 0.0), 120.0, 128, 6.0, 1 / 25.0, seed)


def sorted_keys(d):
    k = d[:, [0, 1, 3]]
    return k[np.lexsort(k.T[::-1])]


def setup(expand):
    spec = small_ruins(seed=71, instances=3000, shadows=True)
    loaded = load_gpu(spec, max_views=3, occlusion=True, max_triangles=20_000_000, max_shadow=10_000_000, expand=expand)
    cams = [ground_camera(*c) for c in CAMERAS]
    views = [make_view(c, 540, slot=i, occlusion=2, error_threshold=[1.0, 2.0, 0.7][i], shadow_sweep=(10.0, -30.0, 5.0))
             for i, c in enumerate(cams)]
    return loaded, cams, views


@pytest.mark.parametrize("expand", [False, True])
def test_two_phases_match_the_oracle(gpu_available, expand):
    loaded, cams, views = setup(expand)
    scene, pool, mm = loaded.scene, loaded.pool, loaded.ctx.config.max_meshlets
    d1 = [depth_for(c, 80 + i) for i, c in enumerate(cams)]
    d2 = [depth_for(c, 90 + i) for i, c in enumerate(cams)]
    for v, d in zip(views, d1):
        scene.set_hzb(v.slot, d)
    refs = [oracle.cut_two_phase(pool, v, oracle.hzb_build(a, 1024), oracle.hzb_build(b, 1024)) for v, a, b in zip(views, d1, d2)]
    print("first phase", flush=True)
    scene.cut(views)
    scene.sync()
    first = [scene.read(v.slot) for v in views]
    idx1 = [scene.read_indices(v.slot, 0) for v in views] if expand else None
    for v, g, r in zip(views, first, refs):
        compare_cut(g, r.first, mm, f"first phase slot {v.slot}")
        n, dropped = scene.phase_rejects(v.slot)
        assert n == r.rejects[1] and not dropped
    assert sum(r.rejects[0] + r.rejects[1] for r in refs) > 0, "the first HZBs must hide something"
    for v, d in zip(views, d2):
        scene.set_hzb(v.slot, d)
    print("second phase", flush=True)
    scene.cut_phase2(views)
    scene.sync()
    added = 0
    for i, (v, r) in enumerate(zip(views, refs)):
        c2, a2 = scene.read_phase2(v.slot)
        diff = np.nonzero(c2 != r.counters2)[0]
        assert len(diff) == 0, f"slot {v.slot}: second-phase counters differ at {diff[:8].tolist()}: {c2[diff[:8]].tolist()} vs {r.counters2[diff[:8]].tolist()}"
        assert np.array_equal(a2, r.args2), f"slot {v.slot}: second-phase args differ"
        draws = scene.read(v.slot)[2]
        n1, n2 = int(first[i][0][2]), min(int(c2[2]), mm)
        assert np.array_equal(sorted_keys(draws[0, :n1]), sorted_keys(r.draws[0, :n1])), f"slot {v.slot}: first-phase set changed"
        assert np.array_equal(sorted_keys(draws[0, n1:n2]), sorted_keys(r.draws[0, n1:n2])), f"slot {v.slot}: second-phase sets differ"
        second = np.zeros((2, mm, 4), dtype=np.uint32)
        second[0, :n2 - n1] = draws[0, n1:n2]
        assert tiling_ok(second[0], n2 - n1, c2, 0, 64, loaded.slot_of_instance), f"slot {v.slot}: second-phase firstIndex tiling"
        added += n2 - n1
        if expand:
            got = scene.read_indices(v.slot, 0)
            want, written = oracle.expand(pool, 0, c2, draws, idx1[i], first=n1)
            assert np.array_equal(got, want), f"slot {v.slot}: second-phase indices differ"
            assert written == int(c2[2 + 4]) - int(first[i][0][2 + 4])
    assert added > 0, "the second HZBs must reveal something"
    with pytest.raises(native.VgError) as e:  # a second phase needs the cut right before it
        scene.cut_phase2(views)
    assert e.value.code == native.VG_E_INVALID


def test_two_phase_properties(gpu_available):
    loaded, cams, views = setup(False)
    scene, mm = loaded.scene, loaded.ctx.config.max_meshlets
    far = np.ones((H, W), dtype=np.float32)
    for v, c in zip(views, cams):
        scene.set_hzb(v.slot, depth_for(c, 80 + v.slot))
    # the cut without occlusion, from the same controller state
    plain_views = [make_view(c, 540, slot=v.slot, occlusion=False, error_threshold=[1.0, 2.0, 0.7][v.slot],
                             shadow_sweep=(10.0, -30.0, 5.0)) for v, c in zip(views, cams)]
    scene.cut(plain_views)
    scene.sync()
    plain = [scene.read(v.slot) for v in views]
    for v in views:
        scene.set_state(v.slot, oracle.initial_state())
    # against the same HZB the second phase adds nothing
    scene.cut(views)
    scene.sync()
    n1 = [int(scene.read(v.slot)[0][2]) for v in views]
    for v, c in zip(views, cams):
        scene.set_hzb(v.slot, depth_for(c, 80 + v.slot))
    scene.cut_phase2(views)
    scene.sync()
    for v, n in zip(views, n1):
        c2, _ = scene.read_phase2(v.slot)
        assert int(c2[2]) == n and not c2[12:12 + 64 * 8:2].any()
    # against an HZB that hides nothing the two phases draw the cut without occlusion
    for v in views:
        scene.set_state(v.slot, oracle.initial_state())
    scene.cut(views)
    scene.sync()
    for v in views:
        scene.set_hzb(v.slot, far)
    scene.cut_phase2(views)
    scene.sync()
    for v, p in zip(views, plain):
        c2, _ = scene.read_phase2(v.slot)
        draws = scene.read(v.slot)[2]
        total = int(c2[2])
        assert total == int(p[0][2])
        assert np.array_equal(sorted_keys(draws[0, :total]), sorted_keys(p[2][0, :total]))
