"""The occluder raster on the GPU (vg_hzb_from_cut) against the oracle: the depth image rasterized from a cut's
camera entries and the HZB built from it equal the oracle's raster (or_raster) and HZB build bit for bit,
for the view's own camera and another one; and a frame loop driven by it, as the reference's occluder pass
drives occlusion (OcclusionCulling.ts:4-21): phase 1 against the raster of the previous frame's cut
(both phases) with the current camera, phase 2 against the raster of phase 1's cut, every phase equal to
the oracle's two-phase cut over the oracle's own rasters."""
import numpy as np
import pytest

from helpers import compare_cut, draw_set, ground_camera, load_gpu, small_ruins
from oracle import oracle
from vgcut import camera as cam
from vgcut import dag, native
from vgcut.virtual_geometry import make_view

pytestmark = pytest.mark.gpu

W, H = 480, 270


def vp_of(camera):
    return np.asarray(cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse), dtype=np.float32)


def oracle_hzb(pool, draws, n, vp):
    depth = oracle.raster(pool, draws, n, vp, W, H)
    return depth, oracle.hzb_build(depth, 1024)


def test_raster_equals_the_oracle(gpu_available):
    spec = small_ruins(seed=31, instances=3000)
    loaded = load_gpu(spec, max_views=2)
    scene, pool = loaded.scene, loaded.pool
    cams = [ground_camera((0.0, 2.0, -70.0), (0.0, 0.0, 60.0)), ground_camera((50.0, 2.0, 10.0), (-60.0, 2.0, 0.0))]
    views = [make_view(c, 540, slot=i) for i, c in enumerate(cams)]
    scene.cut(views)
    scene.sync()
    covered = []
    for v, c in zip(views, cams):
        counters, _, draws, _ = scene.read(v.slot)
        n = min(int(counters[2]), loaded.ctx.config.max_meshlets)
        assert n > 0
        for other in cams:  # the slot's own camera, then the other one (a moved camera, as the next frame's)
            vp = vp_of(other)
            print(f"raster slot {v.slot}: {n} entries", flush=True)
            scene.hzb_from_cut(v.slot, vp, W, H)
            got = scene.hzb(v.slot)
            depth, hzb = oracle_hzb(pool, draws[0], n, vp)
            assert np.array_equal(got[:len(hzb)], hzb), f"slot {v.slot}: raster HZB differs from the oracle"
            covered.append(float((depth < 1).mean()))
    assert max(covered) > 0.05, covered


def test_frame_loop_driven_by_the_raster(gpu_available):
    spec = small_ruins(seed=71, instances=3000)
    loaded = load_gpu(spec, max_views=1, occlusion=True, max_triangles=20_000_000, max_shadow=10_000_000)
    scene, pool, mm = loaded.scene, loaded.pool, loaded.ctx.config.max_meshlets
    path = [((-30.0 + 6 * f, 2.0, -60.0 + 10 * f), (10.0 - 4 * f, 0.5, 50.0)) for f in range(4)]
    state = oracle.initial_state()
    prev_draws, prev_n = None, 0
    hidden = added = 0
    for f, (eye, target) in enumerate(path):
        camera = ground_camera(eye, target)
        vp = vp_of(camera)
        view = make_view(camera, 540, occlusion=2)
        print(f"frame {f}", flush=True)
        if prev_draws is None:
            scene.set_hzb(0, np.ones((H, W), dtype=np.float32))
            hzb1 = oracle.hzb_build(np.ones((H, W), dtype=np.float32), 1024)
        else:
            scene.hzb_from_cut(0, vp, W, H, with_phase2=True)  # last frame's cut, both phases, current camera
            _, hzb1 = oracle_hzb(pool, prev_draws, prev_n, vp)
            assert np.array_equal(scene.hzb(0)[:len(hzb1)], hzb1), f"frame {f}: first HZB differs"
        scene.cut([view])
        scene.sync()
        first = scene.read(0)
        ref1 = oracle.cut(pool, view, state=state, hzb=hzb1)
        compare_cut(first, ref1, mm, f"frame {f} first phase")
        n1 = min(int(ref1.counters[2]), mm)
        scene.hzb_from_cut(0, vp, W, H)  # this frame's first phase
        _, hzb2 = oracle_hzb(pool, ref1.draws[0], n1, vp)
        assert np.array_equal(scene.hzb(0)[:len(hzb2)], hzb2), f"frame {f}: second HZB differs"
        ref = oracle.cut_two_phase(pool, view, hzb1, hzb2, state=state)
        scene.cut_phase2([view])
        scene.sync()
        c2, a2 = scene.read_phase2(0)
        assert np.array_equal(c2, ref.counters2), f"frame {f}: second-phase counters differ"
        assert np.array_equal(a2, ref.args2), f"frame {f}: second-phase args differ"
        draws = scene.read(0)[2]
        n2 = min(int(c2[2]), mm)
        assert np.array_equal(draw_set(draws[0][n1:n2], n2 - n1), draw_set(ref.draws[0][n1:n2], n2 - n1))
        hidden += int(ref1.counters[1])
        added += n2 - n1
        state = ref.first.state
        prev_draws, prev_n = ref.draws[0], n2
    assert hidden > 0, "the rasterized occluders must hide something"


def test_chunk_queue_overflow_and_a_smaller_image(gpu_available):
    """N copies of a quad filling the image: every triangle's pixel box is the whole image, so it is queued
    as (W / 64) x (H / 64) rectangles and the 2^20-rectangle queue overflows. The first raster (1024 x 1024,
    256 per triangle, a multiple of the capacity) fills every queue slot; the second (1024 x 576, 144 per
    triangle) has a reservation running past the capacity, whose slots inside the queue must not replay the
    first raster's rectangles. The second raster is also smaller than the first: the HZB buffer must hold
    none of the first image. Both must equal the oracle, whole buffers compared (one quad's raster: the
    copies coincide)."""
    S, N = 200.0, 3700
    pos = np.array([[-S, -S, 0.0], [S, -S, 0.0], [S, S, 0.0], [-S, S, 0.0]], np.float32)
    data = dag.build_virtual_mesh(pos, np.array([0, 1, 2, 0, 2, 3], np.uint32))
    assert data.meshletCount == 1
    ctx = native.Context(max_views=1, mesh_slots=2, device=0)
    scene = native.Scene(ctx)
    slot = scene.add_mesh(data)
    scene.set_instances(slot, np.tile(np.eye(4, dtype=np.float32).reshape(1, 16), (N, 1)))
    pool = oracle.pool_from_scene(scene, ctx.config.max_meshlets, ctx.config.max_triangles, ctx.config.max_shadow_triangles)
    chunks = []
    for distance, w, h in ((5.0, 1024, 1024), (50.0, 1024, 576)):
        camera = cam.perspective_camera(60, w / h, 0.1, 20000).look_at((0.0, 0.0, distance), (0.0, 0.0, 0.0))
        vp = vp_of(camera)
        scene.cut([make_view(camera, h)])
        scene.sync()
        counters, _, draws, _ = scene.read(0)
        assert int(counters[2]) == N
        chunks.append(2 * N * (w // 64) * (h // 64))
        print(f"quad raster {w} x {h}: {chunks[-1]} rectangles", flush=True)
        scene.hzb_from_cut(0, vp, w, h)
        depth = oracle.raster(pool, draws[0], 1, vp, w, h)
        assert np.all(depth < 1)
        want = oracle.hzb_build(depth, ctx.config.hzb_max_size)
        got = scene.hzb(0)
        assert np.array_equal(got[:len(want)], want), f"{w} x {h}: HZB differs from the oracle"
    capacity = 1 << 20
    assert chunks[0] > capacity and chunks[0] % (2 * 16 * 16) == 0 and capacity % (16 * 16) == 0
    assert chunks[1] > capacity and capacity % (16 * 9) != 0


def test_batched_rasters_equal_single_ones(gpu_available):
    """vg_hzb_from_cuts: three slots of different image sizes and cameras drawn in one launch sequence give
    each slot the HZB that its own vg_hzb_from_cut and the oracle's raster give, whole buffers compared; a
    slot named twice is refused."""
    spec = small_ruins(seed=41, instances=3000)
    loaded = load_gpu(spec, max_views=3)
    scene, pool = loaded.scene, loaded.pool
    cams = [ground_camera((0.0, 2.0, -70.0), (0.0, 0.0, 60.0)), ground_camera((50.0, 2.0, 10.0), (-60.0, 2.0, 0.0)),
            ground_camera((-40.0, 6.0, 30.0), (30.0, 0.0, -20.0))]
    sizes = [(480, 270), (320, 200), (640, 360)]
    views = [make_view(c, 540, slot=i) for i, c in enumerate(cams)]
    scene.cut(views)
    scene.sync()
    rasters = [(v.slot, vp_of(cams[(i + 1) % 3]), w, h, False) for i, (v, (w, h)) in enumerate(zip(views, sizes))]
    print("batched raster of 3 slots", flush=True)
    scene.hzb_from_cuts(rasters)
    batched = [scene.hzb(v.slot) for v in views]
    for (slot, vp, w, h, _), got in zip(rasters, batched):
        counters, _, draws, _ = scene.read(slot)
        n = min(int(counters[2]), loaded.ctx.config.max_meshlets)
        depth = oracle.raster(pool, draws[0], n, vp, w, h)
        want = oracle.hzb_build(depth, loaded.ctx.config.hzb_max_size)
        assert np.array_equal(got[:len(want)], want), f"slot {slot}: batched raster HZB differs from the oracle"
        assert float((depth < 1).mean()) > 0.02
        print(f"single raster of slot {slot}", flush=True)
        scene.hzb_from_cut(slot, vp, w, h)
        assert np.array_equal(scene.hzb(slot), got), f"slot {slot}: batched and single rasters differ"
    with pytest.raises(native.VgError):
        scene.hzb_from_cuts([rasters[0], rasters[0]])
