"""Known answers for the parts of the oracle a GPU-vs-oracle comparison cannot check (a mistake shared by
both sides would pass it): the capacity controller, the prefix and draw-args layout, the distance and
size culls of cells and instances, and the view constants (projection, projScale, frustum planes).

Every expected value below is worked out by hand from the reference's text, not by running the oracle's
formula again:
- controller (GeometryPool.ts:1027-1069): fill = max(requested / capacity, selected / maxDrawn); a cut
  over capacity is redone at scale * clamp(sqrt(fill / 0.82), 1.1, 4), capped at lodScaleLimit; between
  frames the scale is multiplied by min(sqrt(fill / 0.82), 1.03) above fill 0.9 and divided by 1.01 below
  0.75, kept in [1, limit]; words are x10000 rounded to even;
- prefix (GeometryPool.ts:973-995): regions laid out slot-major, part 0 then part 1; args =
  (min(requested, room), 1, start, 0, 0), room = capacity - start (0 past it);
- culls (GeometryPool.ts:742-757, 877-893): placed 1e-4 relative on either side of each boundary, where
  binary32 rounding cannot move a decision;
- view math (VirtualGeometry.ts:414-437, three.js formulas): P and the six planes of axis-aligned
  perspective and orthographic cameras in closed form, projScale = P[5] * 540 = 935.307 at 60 degrees.

Each check is then shown to catch a plausible mistake: the oracle compiled with one planted mistake
(-DOR_MUTANT=k, on the CPU) fails at least one known answer, and so do mistaken view-math functions.
"""
import functools
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle
from vgcut import camera as cam
from vgcut import native, scenegen, scenes
from vgcut.virtual_geometry import make_view

MUTANTS = range(1, 12)


# ------------------------------------------------------------------ capacity controller

def controller_failures(lib=None):
    out = []

    def check(label, steer, requested, selected, want_state, want_redo=None, state=None, limit=64.0, capacity=(10_000, 10_000),
              max_meshlets=1_000):
        st, redo = oracle.controller(steer, requested, selected, capacity, max_meshlets, limit, state=state, lib=lib)
        if not np.array_equal(st, np.asarray(want_state, dtype=np.uint32)):
            out.append(f"{label}: state {st.tolist()} != {list(want_state)}")
        if want_redo is not None and redo != want_redo:
            out.append(f"{label}: redo {redo} != {want_redo}")

    # check: fill 3.28 = 0.82 * 4 -> sqrt(4) = 2: camera redone at scale 2 (retry word 1); shadow fits
    check("retry x2", False, (32_800, 0), (0, 0), (20000, 10000, 0, 0, 10000, 0, 0, 0), [1, 0])
    # fill 1.0001 -> sqrt(1.0001 / 0.82) = 1.104356: above the 1.1 floor, so the target fill shows
    check("retry just over", False, (10_001, 0), (0, 0), (11044, 10000, 0, 0, 10000, 0, 0, 0), [1, 0])
    # fill 100 -> sqrt(121.95) = 11.04, clamped to x4
    check("retry clamped", False, (1_000_000, 0), (0, 0), (40000, 10000, 0, 0, 10000, 0, 0, 0), [1, 0])
    # the limit caps it: x4 from 1 would be 4, limit 3
    check("retry at the limit", False, (1_000_000, 0), (0, 0), (30000, 10000, 0, 0, 10000, 0, 0, 0), [1, 0], limit=3.0)
    # selected meshlets set the fill too (2 x maxDrawn -> sqrt(2 / 0.82) = 1.561738), on the shadow cut
    check("retry by meshlets", False, (0, 0), (5, 2_000), (10000, 15617, 0, 0, 0, 10000, 0, 0), [0, 1])
    # at the limit already: no redo, state as it was
    check("no room to coarsen", False, (40_000, 0), (0, 0), (30000, 10000, 0, 0, 0, 0, 0, 0), [0, 0],
          state=(30000, 10000, 0, 0, 0, 0, 0, 0), limit=3.0)
    # fill exactly 1 is not over capacity
    check("full, not over", False, (10_000, 0), (0, 0), (10000, 10000, 0, 0, 0, 0, 0, 0), [0, 0])
    # steer: any fill above 0.9 gives exactly x1.03 (sqrt(0.9 / 0.82) = 1.0477 > 1.03); fill words 9500, 0
    check("steer up", True, (9_500, 0), (0, 0), (20600, 10000, 9500, 0, 0, 0, 0, 0), state=(20000, 10000, 0, 0, 0, 0, 0, 0))
    check("steer up, heavy", True, (30_000, 0), (0, 0), (10300, 10000, 30000, 0, 0, 0, 0, 0))
    # fill 0.8 is inside the dead band: no change
    check("dead band", True, (8_000, 0), (0, 0), (20000, 10000, 8000, 0, 0, 0, 0, 0), state=(20000, 10000, 0, 0, 0, 0, 0, 0))
    # fill 0.7 -> 2 / 1.01 = 1.980198; the shadow cut at scale 1 stays at 1 (never below)
    check("steer down", True, (7_000, 0), (0, 0), (19802, 10000, 7000, 0, 0, 0, 0, 0), state=(20000, 10000, 0, 0, 0, 0, 0, 0))
    # the steer keeps the scale under the limit (limit 2, fill 0.95: 2 * 1.03 -> 2)
    check("steer capped", True, (9_500, 0), (0, 0), (20000, 10000, 9500, 0, 0, 0, 0, 0), state=(20000, 10000, 0, 0, 0, 0, 0, 0),
          limit=2.0)
    return out


def test_controller_known_answers():
    assert controller_failures() == []


# ------------------------------------------------------------------ prefix and draw args

def prefix_failures(lib=None):
    out = []
    # 3 slots x 2 parts of index counts, capacity 100, regions slot-major:
    #   slot 0: part 0 [0, 30), part 1 [30, 42); slot 1: part 0 [42, 42), part 1 [42, 87);
    #   slot 2: part 0 [87, 111) clamped to the 13 indices left, part 1 [111, 120) past the capacity (0).
    counts = [[30, 12], [0, 45], [24, 9]]
    want_pairs = [30, 0, 12, 30, 0, 42, 45, 42, 24, 87, 9, 111]
    want_args = [30, 1, 0, 0, 0, 12, 1, 30, 0, 0, 0, 1, 42, 0, 0, 45, 1, 42, 0, 0, 13, 1, 87, 0, 0, 0, 1, 111, 0, 0]
    for cut in (0, 1):
        counters, args = oracle.prefix(4, counts, 100, cut=cut, selected=7, max_meshlets=5, lib=lib)
        pairs = counters[12 + cut * 16: 12 + cut * 16 + 12].tolist()
        if pairs != want_pairs:
            out.append(f"cut {cut}: (count, start) pairs {pairs}")
        a = args[cut * 40: cut * 40 + 30].tolist()
        if a != want_args:
            out.append(f"cut {cut}: args {a}")
        if args[cut * 40 + 30:].any() or args[(1 - cut) * 40:(1 - cut) * 40 + 40].any():
            out.append(f"cut {cut}: args outside the used slots were written")
        if counters[2 + cut * 5 + 4] != 120 or counters[2 + cut * 5 + 1] != 5:
            out.append(f"cut {cut}: requested {counters[2 + cut * 5 + 4]} (120), drawn {counters[2 + cut * 5 + 1]} (min(7, 5))")
    # The second phase lays its regions out after the first phase's requested indices (offset 120).
    counters, args = oracle.prefix(4, [[6, 0], [0, 3]], 125, offset=120, lib=lib)
    if counters[12:20].tolist() != [6, 120, 0, 126, 0, 126, 3, 126] or args[:20].tolist() != [
            5, 1, 120, 0, 0, 0, 1, 126, 0, 0, 0, 1, 126, 0, 0, 0, 1, 126, 0, 0]:
        out.append(f"offset 120: pairs {counters[12:20].tolist()} args {args[:20].tolist()}")
    return out


def test_prefix_known_answers():
    assert prefix_failures() == []


# ------------------------------------------------------------------ distance and size culls

@functools.lru_cache(maxsize=None)
def small_sphere():
    return scenegen.icosphere(triangles=2_000, radius=1.0, seed=3, amplitude=0.0)


class CullScene:
    """One small mesh (bounding sphere (c, r)) with instances translated along -z in front of a camera at
    the origin looking down -z (so the view matrix is the identity), on a host-only context."""

    def __init__(self, positions, max_draw=1e30, min_px=0.0):
        self.data = small_sphere()
        mats = np.zeros((len(positions), 16), dtype=np.float32)
        mats[:, 0] = mats[:, 5] = mats[:, 10] = mats[:, 15] = 1.0
        mats[:, 12:15] = positions
        self.ctx = native.Context(max_meshlets=1 << 12, device=-1)
        self.scene = native.Scene(self.ctx)
        slot = self.scene.add_mesh(self.data)
        self.scene.set_instances(slot, mats)
        self.scene.set_flags(slot, max_draw_distance=max_draw, min_pixel_radius=min_px, shadow=False)
        self.pool = oracle.pool_from_scene(self.scene, 1 << 12, 10_000_000, 4_000_000)
        cells = self.pool.cells.reshape(-1, 8).astype(np.float64)
        assert len(cells) == 1
        self.cell_gap = math.sqrt(float((cells[0, :3] ** 2).sum())) - cells[0, 3]
        self.max_r = cells[0, 4]
        b = np.asarray(self.data.boundingSphere, dtype=np.float64)
        self.inst_dist = [math.sqrt(float(((np.asarray(p, dtype=np.float64) + b[:3]) ** 2).sum())) for p in positions]
        self.inst_r = float(b[3])
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 1000).look_at((0, 0, 0), (0, 0, -1))
        self.view = make_view(camera, 1080, min_pixel_radius=0.0)
        self.ps = float(self.view.proj_scale)

    def run(self, lib=None):
        r = oracle.cut(self.pool, self.view, lib=lib)
        return int(r.counters[0]), int(r.stats.instances_tested), int(r.counters[4])  # visible cells, tested, survived

    def close(self):
        self.scene.close()
        self.ctx.close()


def cull_failures(lib=None):
    out = []

    def expect(label, positions, want, max_draw=1e30, min_px=0.0):
        sc = CullScene(positions)
        sc.close()
        got_scene = CullScene(positions, max_draw=max_draw(sc) if callable(max_draw) else max_draw,
                              min_px=min_px(sc) if callable(min_px) else min_px)
        got = got_scene.run(lib)
        got_scene.close()
        if got != want:
            out.append(f"{label}: (visible cells, instances tested, survived) {got} != {want}")

    one = [(0.0, 0.0, -100.0)]
    two = [(0.0, 0.0, -100.0), (60.0, 0.0, -100.0)]  # the cell reaches nearer than either instance
    # maxDrawDistance against the gap (distance minus radius): cell and instance share it with one instance
    # (inst_dist: distance to the instance's sphere centre; the gap is that minus its radius)
    expect("draw distance, one instance, inside", one, (1, 1, 1), max_draw=lambda s: s.inst_dist[0] - s.inst_r + 0.01)
    expect("draw distance, one instance, beyond", one, (0, 0, 0), max_draw=lambda s: s.inst_dist[0] - s.inst_r - 0.01)
    # with two, the cell passes and the nearer instance decides
    expect("draw distance, instance beyond", two, (1, 2, 0), max_draw=lambda s: s.inst_dist[0] - s.inst_r - 0.01)
    expect("draw distance, instance inside", two, (1, 2, 1), max_draw=lambda s: s.inst_dist[0] - s.inst_r + 0.01)
    # minPixelRadius (the mesh's lever; the view's is 0) against r * projScale / distance for the instance
    px = lambda s: s.inst_r * s.ps / s.inst_dist[0]
    expect("min pixels, instance kept", one, (1, 1, 1), min_px=lambda s: px(s) * (1 - 1e-4))
    expect("min pixels, instance culled", one, (1, 1, 0), min_px=lambda s: px(s) * (1 + 1e-4))
    # ... and maxR * projScale / gap for the cell (larger than the instance's: gap < distance)
    cell_px = lambda s: s.max_r * s.ps / s.cell_gap
    expect("min pixels, cell culled", two, (0, 0, 0), min_px=lambda s: cell_px(s) * (1 + 1e-4))
    expect("min pixels, cell kept", two, (1, 2, 0), min_px=lambda s: cell_px(s) * (1 - 1e-4))
    # a cell whose gap is at most maxR is never too small, however large the lever; its instance is
    near = [(0.0, 0.0, -2.6)]  # distance 2.67 > r = 1.73, gap 0.94 <= maxR
    expect("min pixels, cell around the camera", near, (1, 1, 0), min_px=1e6)
    return out


def test_cull_known_answers():
    sc = CullScene([(0.0, 0.0, -100.0)])
    assert abs(sc.ps - 935.3074360871938) < 1e-3
    sc.close()
    assert cull_failures() == []


# ------------------------------------------------------------------ view constants

def view_failures(perspective=cam.perspective, make=make_view):
    out = []
    rel = lambda a, b: abs(a - b) <= 1e-12 * max(1.0, abs(b))
    near, far = 0.1, 1000.0
    camera = cam.Camera(perspective(60, 16 / 9, near, far), near, far, True).look_at((0, 0, 0), (0, 0, -1))
    p = camera.projectionMatrix
    s3 = math.sqrt(3.0)  # cot(30 degrees)
    want = np.zeros(16)
    want[0], want[5], want[10], want[11], want[14] = s3 * 9 / 16, s3, -far / (far - near), -1.0, -far * near / (far - near)
    if not all(rel(float(a), float(b)) for a, b in zip(p, want)):
        out.append(f"perspective P {np.asarray(p).tolist()}")
    v = make(camera, 1080)
    if abs(v.proj_scale - np.float32(s3 * 540)) > np.spacing(np.float32(935.3)):
        out.append(f"projScale {v.proj_scale} != 935.30744")
    planes = np.array(v.frustum_planes[:], dtype=np.float64).reshape(6, 4)
    a, b = s3 * 9 / 16, s3
    ka, kb = math.sqrt(a * a + 1), math.sqrt(b * b + 1)
    want_planes = np.array([[-a / ka, 0, -1 / ka, 0], [a / ka, 0, -1 / ka, 0], [0, b / kb, -1 / kb, 0], [0, -b / kb, -1 / kb, 0],
                            [0, 0, 1, far], [0, 0, -1, -near]])
    if np.abs(planes - want_planes).max() > 1e-6 * far:
        out.append(f"perspective planes {planes.tolist()}")
    ortho = cam.Camera(cam.orthographic(-60, 60, 60, -60, near, 500.0), near, 500.0, False).look_at((0, 0, 0), (0, 0, -1))
    vo = make(ortho, 2048)
    want_planes = np.array([[-1, 0, 0, 60], [1, 0, 0, 60], [0, 1, 0, 60], [0, -1, 0, 60], [0, 0, 1, 500], [0, 0, -1, -near]])
    planes = np.array(vo.frustum_planes[:], dtype=np.float64).reshape(6, 4)
    if np.abs(planes - want_planes).max() > 1e-4:
        out.append(f"orthographic planes {planes.tolist()}")
    # A moved and turned camera: a world point's signed distance to each world plane equals its view-space
    # point's distance to the view-space plane (the view matrix is rigid, so normalisation carries over).
    moved = cam.Camera(perspective(60, 16 / 9, near, far), near, far, True).look_at((30.0, 12.0, -7.0), (-5.0, 2.0, 40.0))
    vm = make(moved, 1080)
    world_planes = np.array(vm.frustum_planes[:], dtype=np.float64).reshape(6, 4)
    V = np.asarray(moved.matrixWorldInverse, dtype=np.float64).reshape(4, 4).T
    rng = np.random.default_rng(5)
    pts = np.c_[rng.uniform(-200, 200, (64, 3)), np.ones(64)]
    d_world = pts @ world_planes.T
    d_view = (pts @ V.T) @ np.array(want_planes_perspective(near, far)).T
    if np.abs(d_world - d_view).max() > 1e-3:
        out.append("moved camera: world planes do not follow the view matrix")
    return out


def want_planes_perspective(near, far):
    s3 = math.sqrt(3.0)
    a, b = s3 * 9 / 16, s3
    ka, kb = math.sqrt(a * a + 1), math.sqrt(b * b + 1)
    return [[-a / ka, 0, -1 / ka, 0], [a / ka, 0, -1 / ka, 0], [0, b / kb, -1 / kb, 0], [0, -b / kb, -1 / kb, 0], [0, 0, 1, far],
            [0, 0, -1, -near]]


def test_view_constants_closed_form():
    assert view_failures() == []


def test_view_math_mistakes_are_caught(monkeypatch):
    # the full field of view taken as the half angle
    assert view_failures(perspective=lambda fov, aspect, n, f: cam.perspective(2 * fov, aspect, n, f))
    # planes extracted from V * P instead of P * V
    real_mul = cam.mat_mul
    monkeypatch.setattr(cam, "mat_mul", lambda a, b: real_mul(b, a))
    assert view_failures()
    monkeypatch.setattr(cam, "mat_mul", real_mul)
    # planes left unnormalised
    real_planes = cam.frustum_planes
    monkeypatch.setattr(cam, "frustum_planes", lambda m: real_planes(m) * 2.0)
    assert view_failures()


# ------------------------------------------------------------------ planted mistakes in the oracle

def all_failures(lib):
    return controller_failures(lib) + prefix_failures(lib) + cull_failures(lib)


def test_planted_oracle_mistakes_fail_the_known_answers(tmp_path):
    paths = [str(tmp_path / f"liboracle_m{k}.so") for k in MUTANTS]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda a: oracle.build_variant(*a), zip(paths, MUTANTS)))
    assert all_failures(None) == []
    for k, path in zip(MUTANTS, paths):
        assert all_failures(oracle.load_variant(path)), f"the oracle with mistake {k} passes every known answer"
