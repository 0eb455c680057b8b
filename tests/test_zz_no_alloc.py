"""vg_cut_static, the cut whose launch buffers live in the library's static scratch (it allocates nothing),
against vg_cut, the entry point it is a variant of, through the C ABI: counters (controller state
included), draw args, drawn sets and work metrics bit-exact, and both equal to the oracle. Two calls in a
row share the device's scratch, the second with a smaller input (fewer views), so a stale queue, work
list or counter from the first would show. Sizes past the scratch's limits are refused with a message.
"""
import numpy as np
import pytest

from helpers import compare_cut, ground_camera, load_gpu, small_ruins
from oracle import oracle
from vgcut import camera as cam
from vgcut import native
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu


def views_of(count):
    rng = np.random.default_rng(83)
    out = []
    for slot in range(count):
        if slot % 4 == 3:
            ortho = cam.orthographic_camera(-60, 60, 60, -60, 0.1, 500).look_at((30.0, 200.0, 20.0 + slot), (0.0, 0.0, 0.0))
            out.append(make_view(ortho, 2048, slot=slot))
        else:
            camera = ground_camera((rng.uniform(-50, 50), rng.uniform(2, 40), rng.uniform(-50, 50)),
                                   (rng.uniform(-50, 50), 0.0, rng.uniform(-50, 50)))
            out.append(make_view(camera, [720, 1080, 2160][slot % 3], slot=slot, error_threshold=[1.0, 0.5, 2.0][slot % 3],
                                 shadow_sweep=(20.0, -40.0, 0.0), triangle_budget=300_000 if slot == 1 else 0))
    return out


def test_cut_static_equals_cut(gpu_available):
    spec = small_ruins(seed=83, shadows=True)
    loaded = load_gpu(spec, max_views=6, max_meshlets=1 << 18)
    scene, mm = loaded.scene, loaded.ctx.config.max_meshlets
    views = views_of(6)
    print("no-alloc: 6 views, then 2", flush=True)
    # vg_cut first: it uploads the scene and gives every slot its outputs; its results are the reference.
    scene.cut(views)
    scene.sync()
    plain = {v.slot: scene.read(v.slot) for v in views}
    for first, count in ((0, 6), (2, 2)):
        subset = views[first:first + count]
        for v in subset:
            scene.set_state(v.slot, oracle.initial_state())
        scene.cut_static(subset)
        for v in subset:
            got = scene.read(v.slot)
            if v.triangle_budget:
                _, _, chosen = oracle.budget_pick(loaded.pool, v, v.triangle_budget,
                                                  native.budget_thresholds(0.5, 3.0, loaded.ctx.config.budget_bins))
                ref = oracle.cut(loaded.pool, chosen)
            else:
                ref = oracle.cut(loaded.pool, v)
            compare_cut(got, ref, mm, f"static call of {count} views, slot {v.slot}")
            compare_cut(plain[v.slot], ref, mm, f"vg_cut slot {v.slot}")
            assert np.array_equal(got[0], plain[v.slot][0]) and np.array_equal(got[1], plain[v.slot][1])


def test_cut_static_refuses_what_does_not_fit(gpu_available):
    spec = small_ruins(seed=84, instances=600, meshes=3)
    loaded = load_gpu(spec, max_views=66)
    scene = loaded.scene
    views = [make_view(ground_camera((0.0, 5.0, -40.0), (0.0, 0.0, 0.0)), 1080, slot=k) for k in range(65)]
    with pytest.raises(native.VgError) as e:
        scene.cut_static(views)
    assert e.value.code == native.VG_E_CAPACITY and "64 views" in str(e.value)
    # A scene not yet on the device, a slot that no vg_cut has given outputs yet, a scene change not yet
    # uploaded: each refused, nothing launched.
    with pytest.raises(native.VgError) as e:
        scene.cut_static(views[:1])
    assert e.value.code == native.VG_E_INVALID and "not on the device" in str(e.value)
    scene.cut(views[:1])
    scene.sync()
    with pytest.raises(native.VgError) as e:
        scene.cut_static(views[1:2])
    assert e.value.code == native.VG_E_INVALID and "outputs" in str(e.value)
    scene.cut_static(views[:1])
    scene.set_flags(0, max_draw_distance=50.0)
    with pytest.raises(native.VgError) as e:
        scene.cut_static(views[:1])
    assert e.value.code == native.VG_E_INVALID and "not on the device" in str(e.value)
