"""The DAG builder (csrc/vgdag.cpp): the reference's own builder invariants (tests/preprocess.test.ts:8-94)
on its test meshes, the reference's cut checks (verifyCutCoverage / verifyCutOverlap, cut.ts:83-122) at
many distances, hand-derived partitions (partition.ts:16-97), closed forms and structural properties."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import cut64
from vgcut import dag, scenegen

ERROR_INFINITY = 1e30
MAX_MESHLET_TRIANGLES = 128


def icosahedron_sphere(detail: int):
    """THREE.IcosahedronGeometry(1, detail) welded: each face split into (detail + 1)^2 triangles."""
    t = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    F = [0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3,
         4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1]
    n = detail + 1
    ids, pts, idx = {}, [], []

    def vid(p):
        p = p / np.linalg.norm(p)
        key = tuple(np.round(p, 9))
        if key not in ids:
            ids[key] = len(pts)
            pts.append(p)
        return ids[key]

    for f in range(20):
        a, b, c = V[F[3 * f]], V[F[3 * f + 1]], V[F[3 * f + 2]]
        g = {(i, j): vid(a + (b - a) * i / n + (c - a) * j / n) for i in range(n + 1) for j in range(n + 1 - i)}
        for i in range(n):
            for j in range(n - i):
                idx += [g[i, j], g[i + 1, j], g[i, j + 1]]
                if j < n - i - 1:
                    idx += [g[i + 1, j], g[i + 1, j + 1], g[i, j + 1]]
    return np.array(pts), np.array(idx, dtype=np.uint32)


def noisy_sphere():
    """preprocess.test.ts:55-65: r = 1 + 0.1 sin(9x) cos(7y) on the detail-40 icosahedron."""
    P, idx = icosahedron_sphere(40)
    n = 1 + 0.1 * np.sin(P[:, 0] * 9) * np.cos(P[:, 1] * 7)
    return (P * n[:, None]).astype(np.float32), idx


def plane(size=100.0, segments=64, height=None):
    """THREE.PlaneGeometry(size, size, segments, segments).rotateX(-pi/2) (preprocess.test.ts:67-72)."""
    xs = np.linspace(-size / 2, size / 2, segments + 1)
    gx, gz = np.meshgrid(xs, xs)
    y = height(gx, gz) if height else np.zeros_like(gx)
    pos = np.stack([gx, y, gz], -1).reshape(-1, 3).astype(np.float32)
    w = segments + 1
    idx = []
    for iy in range(segments):
        for ix in range(segments):
            a, b, c, d = ix + w * iy, ix + w * (iy + 1), ix + 1 + w * (iy + 1), ix + 1 + w * iy
            idx += [a, b, d, b, c, d]
    return pos, np.array(idx, dtype=np.uint32)


def terrain():
    return plane(height=lambda x, z: 4 * np.sin(x * 0.2) * np.cos(z * 0.13))


def check_dag(name, d, positions, indices):
    """preprocess.test.ts:13-51 checkDag."""
    info = d.meshletInfo.reshape(-1, 4)
    b = d.meshletBounds.reshape(-1, 16)
    assert d.meshletCount > 0, f"{name}: produces meshlets"
    assert d.lodLevels > 1, f"{name}: builds more than one LOD level"
    assert d.leafTriangles == len(indices) // 3, f"{name}: leaf level covers every input triangle"
    assert np.all(b[:, 9] >= b[:, 8]), f"{name}: parent error >= own error for every meshlet"
    assert np.all((info[:, 0] > 0) & (info[:, 0] <= MAX_MESHLET_TRIANGLES)), f"{name}: 1..128 triangles"
    assert int(np.sum(b[:, 9] >= ERROR_INFINITY)) == d.rootMeshlets, f"{name}: root count matches stats"
    for i in range(d.meshletCount):
        tris, first, _, voff = (int(x) for x in info[i])
        packed = d.meshletTriangles[first // 3:first // 3 + tris]
        local = np.stack([packed & 255, (packed >> 8) & 255, (packed >> 16) & 255], 1).reshape(-1)
        assert np.array_equal(d.meshletVertices[voff + local], d.indices[first:first + 3 * tris]), \
            f"{name}: packed meshlet triangles decode to the original index list"


def check_cuts(name, d, distances, thresholds=(0.5, 1.0, 4.0)):
    """Complete, non-overlapping cuts (cut.ts:83-122) from a camera at each distance."""
    for dist in distances:
        for t in thresholds:
            vm = np.eye(4)
            vm[3, 2] = -dist
            view = {"viewMatrix": vm.reshape(-1), "projScale": 1000.0, "near": 0.1}
            sel = cut64.select_cut(d, view, t)
            assert not cut64.verify_cut_coverage(d, sel), f"{name}: hole at {dist}, {t} px"
            assert not cut64.verify_cut_overlap(d, sel), f"{name}: overlap at {dist}, {t} px"


def enclosed(inner, outer, rel=1e-5):
    d = np.linalg.norm(inner[..., :3] - outer[..., :3], axis=-1)
    return d + inner[..., 3] <= outer[..., 3] * (1 + rel) + 1e-6


@pytest.fixture(scope="module")
def sphere_dag():
    pos, idx = noisy_sphere()
    return pos, idx, dag.build_virtual_mesh(pos, idx)


def test_reference_builder_invariants(sphere_dag):
    """preprocess.test.ts:74-76: noisy sphere, terrain, terrain with prune."""
    pos, idx, d = sphere_dag
    check_dag("noisy sphere", d, pos, idx)
    pos, idx = terrain()
    check_dag("terrain", dag.build_virtual_mesh(pos, idx), pos, idx)
    check_dag("terrain (prune)", dag.build_virtual_mesh(pos, idx, prune=True), pos, idx)


def test_complete_cuts_at_every_distance(sphere_dag):
    _, _, d = sphere_dag
    check_cuts("noisy sphere", d, (1.2, 2, 5, 20, 60, 200, 800, 5000, 1e5))
    pos, idx = terrain()
    check_cuts("terrain", dag.build_virtual_mesh(pos, idx), (5, 20, 60, 200, 800, 5000))


def test_leaves_partition_the_input(sphere_dag):
    """The leaf meshlets hold exactly the input triangles (as corner triples, in any order)."""
    _, idx, d = sphere_dag
    info = d.meshletInfo.reshape(-1, 4)
    leaf = info[info[:, 2] == 0]
    got = np.concatenate([d.indices[f:f + 3 * t] for t, f, _, _ in leaf]).reshape(-1, 3)
    want = idx.reshape(-1, 3)
    key = lambda a: np.sort(a[:, 0].astype(np.int64) << 42 | a[:, 1].astype(np.int64) << 21 | a[:, 2])  # noqa: E731
    assert np.array_equal(key(got), key(want))
    assert np.all(d.indices < len(d.positions) // 3)
    counts = np.diff(np.append(info[:, 3], len(d.meshletVertices)))
    assert np.all(counts <= 128)


def test_bound_spheres_enclose(sphere_dag):
    """Parent spheres enclose the LOD spheres of their children; the mesh's bound spheres enclose all."""
    _, _, d = sphere_dag
    b = d.meshletBounds.reshape(-1, 16).astype(np.float64)
    non_root = b[:, 9] < ERROR_INFINITY
    assert np.all(enclosed(b[non_root, 0:4], b[non_root, 4:8]))
    assert np.all(enclosed(b[:, 0:4], d.lodBoundsSphere.astype(np.float64)))
    assert np.all(enclosed(b[non_root, 4:8], d.lodBoundsSphere.astype(np.float64)))
    assert np.all(enclosed(b[:, 12:16], d.cullBoundsSphere.astype(np.float64)))
    # Each culling sphere holds its meshlet's vertices.
    info = d.meshletInfo.reshape(-1, 4)
    P = d.positions.reshape(-1, 3).astype(np.float64)
    for i in range(0, d.meshletCount, 7):
        t, f, _, _ = info[i]
        v = P[d.indices[f:f + 3 * t]]
        assert np.all(np.linalg.norm(v - b[i, 12:15], axis=1) <= b[i, 15] * (1 + 1e-6) + 1e-6)


def test_levels_are_contiguous_and_tables_match(sphere_dag):
    """pack() (buildVirtualMesh.ts:484-601): level ranges, level errors and guard errors from the meshlets."""
    _, _, d = sphere_dag
    info = d.meshletInfo.reshape(-1, 4)
    b = d.meshletBounds.reshape(-1, 16)
    assert np.all(np.diff(info[:, 2].astype(np.int64)) >= 0)
    for level in range(len(d.levelRanges) // 2):
        s, c = int(d.levelRanges[2 * level]), int(d.levelRanges[2 * level + 1])
        assert c > 0 and np.all(info[s:s + c, 2] == level)
        assert d.levelErrors[2 * level] == b[s:s + c, 8].min() and d.levelErrors[2 * level + 1] == b[s:s + c, 9].max()
        assert d.levelGuardErrors[2 * level] == b[s:s + c, 8].max()
        assert d.levelGuardErrors[2 * level + 1] == b[s:s + c, 9].min()
    # Replacement lists point one or more levels up, and every child of a list's group shares it.
    g, start, lst = d.replacementGroup, d.replacementStart, d.replacementIndices
    for i in range(d.meshletCount):
        if g[i] >= 0:
            r = lst[start[g[i]]:start[g[i] + 1]]
            assert len(r) > 0 and np.all(info[r, 2] > info[i, 2])
            assert np.all(b[r, 8] == b[i, 9])  # the group's error is its outputs' own error


def test_flat_grid_errors_accumulate_the_floor():
    """A flat grid simplifies without error: each level adds exactly max(shape, 1e-7) = 1e-7
    (buildVirtualMesh.ts:262-270) while group borders are locked, so level l's own error is l * 1e-7."""
    pos, idx = plane(size=10.0, segments=64)
    d = dag.build_virtual_mesh(pos, idx, tailSimplify=False)
    info = d.meshletInfo.reshape(-1, 4)
    b = d.meshletBounds.reshape(-1, 16)
    for level in range(1, len(d.levelRanges) // 2):
        own = b[info[:, 2] == level, 8].astype(np.float64)
        expect = np.float32(0)
        for _ in range(level):
            expect = np.float32(np.float64(expect) + 1e-7)
        assert np.allclose(own, float(expect), rtol=1e-6), (level, own[:4])
    assert d.leafTriangles == 64 * 64 * 2 and d.lodLevels > 2


def test_halves_per_level(sphere_dag):
    """simplifyRatio 0.5: each of the first levels keeps about half the triangles of the one before."""
    _, _, d = sphere_dag
    info = d.meshletInfo.reshape(-1, 4)
    tris = [int(info[info[:, 2] == l, 0].sum()) for l in range(4)]
    for a, c in zip(tris, tris[1:]):
        assert 0.45 * a <= c <= 0.56 * a, tris


def test_deterministic_across_thread_counts(sphere_dag):
    pos, idx, d = sphere_dag
    d1 = dag.build_virtual_mesh(pos, idx, threads=1)
    for f in ("meshletBounds", "meshletInfo", "indices", "meshletVertices", "meshletTriangles", "replacementGroup",
              "replacementStart", "replacementIndices", "levelRanges", "levelErrors", "levelGuardErrors"):
        assert np.array_equal(getattr(d, f), getattr(d1, f)), f


def test_partition_reference_case():
    """preprocess.test.ts:86-92, and the groups it must give: the chain grows from meshlet 0 in Morton
    (x) order, 8 consecutive meshlets at a time."""
    groups = dag.partition_meshlets([{"boundaryEdges": [i, i + 1], "center": [i, 0, 0]} for i in range(40)], 8)
    assert sorted(x for g in groups for x in g) == list(range(40))
    assert all(len(g) <= 8 for g in groups)
    assert groups == [list(range(k, k + 8)) for k in range(0, 40, 8)]


def test_partition_hand_cases():
    """partition.ts:45-96 by hand: score first, then distance to the seed, then undersized groups merged."""
    # 0 shares two edges with 1 and one with 2: the group takes 1 first; 2 is left alone and merged as pending.
    m = [{"boundaryEdges": [1, 2, 3], "center": [0, 0, 0]}, {"boundaryEdges": [1, 2], "center": [5, 0, 0]},
         {"boundaryEdges": [3], "center": [1, 0, 0]}]
    assert dag.partition_meshlets(m, 2) == [[0, 1], [2]]
    # Equal scores: the nearer neighbour (2) joins first.
    m = [{"boundaryEdges": [1, 2], "center": [0, 0, 0]}, {"boundaryEdges": [1], "center": [9, 0, 0]},
         {"boundaryEdges": [2], "center": [1, 0, 0]}]
    assert dag.partition_meshlets(m, 2) == [[0, 2], [1]]
    # Ten isolated meshlets: singletons are undersized (min size max(2, 4 // 2) = 2) and packed 4 at a time.
    m = [{"boundaryEdges": [100 + i], "center": [i, 0, 0]} for i in range(10)]
    assert dag.partition_meshlets(m, 4) == [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9]]
    # No more meshlets than the group size: one group in input order.
    assert dag.partition_meshlets(m[:4], 4) == [[0, 1, 2, 3]]


@pytest.mark.parametrize("kind,size", [(scenegen.SPHERE, (1.0, 0, 0)), (scenegen.BOX, (1.0, 0.6, 0.4)),
                                       (scenegen.COLUMN, (0.4, 4.0, 0)), (scenegen.ARCH, (4.0, 3.0, 0.8))])
def test_procedural_meshes_are_watertight(kind, size):
    """Every edge of the tessellated props has exactly two triangles, once in each direction, and the
    signed volume is positive (outward windings)."""
    pos, idx = scenegen.triangle_mesh(kind, seed=3, size=size, noise_amplitude=0.05, noise_frequency=2.0,
                                      target_triangles=20000)
    t = idx.reshape(-1, 3).astype(np.int64)
    directed = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    keys = directed[:, 0] << 32 | directed[:, 1]
    assert len(np.unique(keys)) == len(keys)
    rev = directed[:, 1] << 32 | directed[:, 0]
    assert np.array_equal(np.sort(keys), np.sort(rev))
    P = pos.astype(np.float64)
    vol = np.einsum("ij,ij->i", P[t[:, 0]], np.cross(P[t[:, 1]], P[t[:, 2]])).sum() / 6
    assert vol > 0
    d = dag.build_virtual_mesh(pos, idx)
    check_dag(f"kind {kind}", d, pos, idx)
    R = float(d.boundingSphere[3])
    check_cuts(f"kind {kind}", d, [R * k for k in (1.5, 4, 16, 64, 256)], thresholds=(1.0,))


def test_invalid_input_is_rejected():
    pos = np.zeros((3, 3), dtype=np.float32)
    with pytest.raises(ValueError):
        dag.build_virtual_mesh(pos, np.array([0, 1, 3], dtype=np.uint32))
    with pytest.raises(ValueError):
        dag.build_virtual_mesh(pos, np.array([0, 1], dtype=np.uint32))
    with pytest.raises(ValueError):
        dag.build_virtual_mesh(pos, np.array([0, 1, 2], dtype=np.uint32), voxelLods=True, voxelResolution=0)
    with pytest.raises(ValueError):
        dag.build_virtual_mesh(pos, np.array([0, 1, 2], dtype=np.uint32), normals=np.zeros(6, np.float32))
    with pytest.raises(TypeError):
        dag.build_virtual_mesh(pos, np.array([0, 1, 2], dtype=np.uint32), voxelLod=True)
    assert math.isfinite(float(dag.build_virtual_mesh(pos + np.eye(3, dtype=np.float32), np.array([0, 1, 2], dtype=np.uint32)).boundingSphere[3]))
