"""Pins the oracle's occluder raster (or_raster, the rules of vg_hzb_from_cut) without retyping it.

- known answers: an axis-aligned quad covers exactly the pixel centres inside it, at its depth;
- exact coverage: triangles whose screen vertices are exact multiples of 1/256 pixel, against an
  independent point-in-triangle test and barycentric depth in rational arithmetic (fractions), both
  windings mixed, the image the minimum over the front faces;
- closed form: a plane facing a perspective camera has depth f (d - n) / (d (f - n)) everywhere;
- rules: back faces, triangles with a vertex behind the camera and meshlets without positions draw nothing.
"""
from fractions import Fraction

import numpy as np

from oracle import oracle

W = H = 16


def pool_of(positions, triangles, instances=None, no_positions=False):
    """A one-meshlet pool: `positions` (n, 3), `triangles` (m, 3) local ids."""
    pos = np.asarray(positions, dtype=np.float32).reshape(-1, 3)
    tri = np.asarray(triangles, dtype=np.uint32).reshape(-1, 3)
    vertices = np.zeros((len(pos), 4), dtype=np.float32)
    vertices[:, :3] = pos
    if instances is None:
        instances = np.zeros(20, dtype=np.float32)
        instances[:16] = np.eye(4, dtype=np.float32).reshape(16)
        instances[16:19] = 1
    info = np.array([len(tri), 0, 0, 0xFFFFFFFF if no_positions else 0], dtype=np.uint32)
    packed = (tri[:, 0] | (tri[:, 1] << 8) | (tri[:, 2] << 16)).astype(np.uint32)
    pool = oracle.Pool(mesh_table=np.zeros(24, np.float32), instances=np.asarray(instances, np.float32).reshape(-1),
                       cells=np.zeros(8, np.float32), bounds=np.zeros(16, np.float32), info=info, slots=1, mesh_slots=1,
                       max_meshlets=1, cut_capacity=(0, 0), triangles=packed, vertices=vertices.reshape(-1),
                       meshlet_vertices=np.arange(len(pos), dtype=np.uint32))
    draws = np.array([[0, 0, 0, len(tri) | (255 << 16)]], dtype=np.uint32)
    return pool, draws


def test_quad_known_answer():
    # Identity view_proj: clip = world, so NDC [-0.5, 0.5]^2 maps to screen [4, 12]^2 on a 16 x 16 image.
    quad = [(-0.5, -0.5, 0.25), (0.5, -0.5, 0.25), (0.5, 0.5, 0.25), (-0.5, 0.5, 0.25)]
    pool, draws = pool_of(quad, [(0, 1, 2), (0, 2, 3)])
    depth = oracle.raster(pool, draws, 1, np.eye(4), W, H)
    want = np.ones((H, W), dtype=np.float32)
    want[4:12, 4:12] = 0.25  # centres 4.5 .. 11.5 lie in [4, 12]; 3.5 and 12.5 do not
    assert np.array_equal(depth, want)
    # Edges through pixel centres count as inside: x in [4.5, 11.5] screen (NDC -0.4375 .. 0.4375).
    quad = [(-0.4375, -0.5, 0.5), (0.4375, -0.5, 0.5), (0.4375, 0.5, 0.5), (-0.4375, 0.5, 0.5)]
    pool, draws = pool_of(quad, [(0, 1, 2), (0, 2, 3)])
    depth = oracle.raster(pool, draws, 1, np.eye(4), W, H)
    want = np.ones((H, W), dtype=np.float32)
    want[4:12, 4:12] = 0.5
    assert np.array_equal(depth, want)


def test_back_faces_and_missing_positions_draw_nothing():
    quad = [(-0.5, -0.5, 0.25), (0.5, -0.5, 0.25), (0.5, 0.5, 0.25), (-0.5, 0.5, 0.25)]
    pool, draws = pool_of(quad, [(0, 2, 1), (0, 3, 2)])  # clockwise in clip space: back faces
    assert np.all(oracle.raster(pool, draws, 1, np.eye(4), W, H) == 1)
    pool, draws = pool_of(quad, [(0, 1, 2), (0, 2, 3)], no_positions=True)
    assert np.all(oracle.raster(pool, draws, 1, np.eye(4), W, H) == 1)
    pool, draws = pool_of(quad, [(0, 1, 2), (0, 2, 3)])
    assert np.all(oracle.raster(pool, draws, 0, np.eye(4), W, H) == 1)  # no entries


def test_vertex_behind_the_camera_drops_its_triangle():
    # Clip z < 0 (in front of the near plane) on vertex 2 only: the first triangle is skipped, the second drawn.
    quad = [(-0.5, -0.5, 0.25), (0.5, -0.5, 0.25), (0.5, 0.5, -0.1), (-0.5, 0.5, 0.25), (-0.5, 0.5, 0.25)]
    pool, draws = pool_of(quad, [(0, 1, 2), (0, 1, 4)])
    depth = oracle.raster(pool, draws, 1, np.eye(4), W, H)
    # (0, 1, 4) is the lower-left half: (-0.5,-0.5), (0.5,-0.5), (-0.5,0.5).
    drawn = np.argwhere(depth < 1)
    assert len(drawn) > 0 and np.all(depth[depth < 1] == np.float32(0.25))
    assert depth[4, 11] == 1.0  # the upper-right corner belongs only to the dropped triangle


def exact_image(tris_screen, zs):
    """Min depth over front-facing triangles, each given by exact screen vertices (Fractions, y down)."""
    img = [[None] * W for _ in range(H)]
    for (a, b, c), (za, zb, zc) in zip(tris_screen, zs):
        def cross(o, p, q):
            return (p[0] - o[0]) * (q[1] - o[1]) - (p[1] - o[1]) * (q[0] - o[0])
        s = cross(a, b, c)
        if s >= 0:  # y down: counter-clockwise in clip space means negative here
            continue
        for j in range(H):
            for i in range(W):
                p = (Fraction(2 * i + 1, 2), Fraction(2 * j + 1, 2))
                la, lb, lc = cross(b, c, p), cross(c, a, p), cross(a, b, p)
                if la <= 0 and lb <= 0 and lc <= 0:
                    z = (la * Fraction(za) + lb * Fraction(zb) + lc * Fraction(zc)) / s
                    z = min(max(z, Fraction(0)), Fraction(1))
                    if img[j][i] is None or z < img[j][i]:
                        img[j][i] = z
    return img


def test_coverage_and_depth_against_rational_arithmetic():
    rng = np.random.default_rng(5)
    # Screen position X / 256 comes from world x = X / 2048 - 1 exactly (identity view_proj, 16 px wide):
    # ((x / 2 + 1 / 2) * 16) * 256 = X with no rounding, so the snapped vertices are known exactly.
    for trial in range(6):
        n = 12
        Xs = rng.integers(-3 * 256, 19 * 256, (n, 3))
        Ys = rng.integers(-3 * 256, 19 * 256, (n, 3))
        Zs = rng.integers(0, 1 << 20, (n, 3)) / float(1 << 20)
        pos, tri = [], []
        for t in range(n):
            for k in range(3):
                pos.append((Xs[t, k] / 2048.0 - 1.0, 1.0 - Ys[t, k] / 2048.0, Zs[t, k]))
            tri.append((3 * t, 3 * t + 1, 3 * t + 2))
        pool, draws = pool_of(pos, tri)
        got = oracle.raster(pool, draws, 1, np.eye(4), W, H)
        screen = [tuple((Fraction(int(Xs[t, k]), 256), Fraction(int(Ys[t, k]), 256)) for k in range(3)) for t in range(n)]
        want = exact_image(screen, [tuple(float(z) for z in Zs[t]) for t in range(n)])
        fronts = 0
        for j in range(H):
            for i in range(W):
                if want[j][i] is None:
                    assert got[j, i] == 1.0, (trial, i, j)
                else:
                    fronts += 1
                    assert abs(float(want[j][i]) - float(got[j, i])) <= 2e-7, (trial, i, j, float(want[j][i]), got[j, i])
        assert fronts > 0


def webgpu_perspective(fov_deg, aspect, near, far):
    """Column-major perspective with WebGPU depth (z_ndc in [0, 1]), as three.js builds it."""
    f = 1.0 / np.tan(np.radians(fov_deg) / 2)
    m = np.zeros((4, 4))
    m[0, 0], m[1, 1] = f / aspect, f
    m[2, 2], m[2, 3] = -far / (far - near), -far * near / (far - near)
    m[3, 2] = -1.0
    return m.T.reshape(16).astype(np.float32)  # column-major


def test_plane_depth_closed_form():
    near, far, d = 0.1, 100.0, 5.0
    vp = webgpu_perspective(60.0, 1.0, near, far)
    big = 20.0  # the quad overfills the view
    quad = [(-big, -big, -d), (big, -big, -d), (big, big, -d), (-big, big, -d)]
    pool, draws = pool_of(quad, [(0, 1, 2), (0, 2, 3)])
    depth = oracle.raster(pool, draws, 1, vp, W, H)
    want = far * (d - near) / (d * (far - near))
    assert np.all(np.abs(depth.astype(np.float64) - want) <= 1e-6)
