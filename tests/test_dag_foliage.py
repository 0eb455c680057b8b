"""The DAG builder's levels for geometry made of separate pieces (csrc/vgdag.cpp): piece-by-piece aggregate
simplification (aggregate.ts) and voxel-proxy tail levels (voxelProxy.ts), both on by default as in the reference
(buildVirtualMesh.ts:161-173).

Restated reference tests: the leaf-card case (tests/preprocess.test.ts:76-86 on tests/leafCards.ts), the voxel
proxy's UVs (tests/import.test.ts:63-70), a torus build's UVs and root count (tests/import.test.ts:53-57), voxel
levels on a conifer-like crown (tests/serialize.test.ts:68-71). Closed forms: pair thinning of a row of equal
squares (error, triangle count and the widened copies' positions), the proxy's closedness and its distance from
the source surface (within VOXEL_ERROR_CELLS cells, voxelProxy.ts:10-15). The recorded-run replay the builder
uses for proxy simplification is checked against fresh runs by tests/dag_replay_check.cpp.
"""
from __future__ import annotations

import collections
import json
import math
import os
import subprocess

import numpy as np
import pytest

import foliage
from oracle import cut64
from test_dag import check_dag
from vgcut import dag

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ERROR_INFINITY = 1e30


@pytest.fixture(scope="module")
def leaves():
    pos, uv, idx, nrm = foliage.leaf_cards()
    return pos, idx, dag.build_virtual_mesh(pos, idx, normals=nrm, uvs=uv)


def test_leaf_cards_reference_case(leaves):
    """preprocess.test.ts:76-86: invariants, groups simplified piece by piece, the first level at most 0.6 of the
    leaf triangles, complete cuts from 5 to 800 m."""
    pos, idx, d = leaves
    check_dag("leaf cards", d, pos, idx)
    assert d.aggregateGroups > 0, "leaf cards: groups are simplified piece by piece"
    info = d.meshletInfo.reshape(-1, 4)
    first, count = (int(x) for x in d.levelRanges[2:4])
    level1 = int(info[first:first + count, 0].sum())
    assert level1 <= d.leafTriangles * 0.6, f"leaf cards: the first level halves the triangles ({level1} of {d.leafTriangles})"
    for distance in (5, 20, 60, 200, 800):
        vm = np.eye(4)
        vm[3, 1], vm[3, 2] = -8, -distance  # makeTranslation(0, -8, -distance).elements (column-major)
        sel = cut64.select_cut(d, {"viewMatrix": vm.reshape(-1), "projScale": 1000.0, "near": 0.1}, 1.0)
        assert not cut64.verify_cut_coverage(d, sel), f"leaf cards: complete cut at {distance} m"
        assert not cut64.verify_cut_overlap(d, sel), f"leaf cards: no overlap at {distance} m"


def test_leaf_cards_attributes_cover_added_vertices(leaves):
    """Thinned pieces and voxel levels append vertices (buildVirtualMesh.ts:388-407): normals and uvs grow with the
    positions, copies keep their source's attributes, and every meshlet vertex id is in range."""
    pos, _, d = leaves
    nv = len(d.positions) // 3
    assert nv > len(pos) // 3, "levels added vertices"
    assert np.array_equal(d.positions[:len(pos)], pos), "the input vertices keep their place"
    assert len(d.normals) == nv * 3 and len(d.uvs) == nv * 2
    assert np.all(np.isfinite(d.normals)) and np.all(np.isfinite(d.uvs))
    assert int(d.indices.max()) < nv and int(d.meshletVertices.max()) < nv
    n = d.normals.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-5), "unit normals"


def test_leaf_cards_same_for_any_thread_count(leaves):
    pos, uv, idx, nrm = foliage.leaf_cards()
    d1 = dag.build_virtual_mesh(pos, idx, normals=nrm, uvs=uv, threads=1)
    _, _, d = leaves
    for f in ("meshletBounds", "meshletInfo", "indices", "positions", "normals", "uvs", "meshletVertices", "meshletTriangles"):
        assert np.array_equal(getattr(d1, f), getattr(d, f)), f


def test_switching_the_levels_off():
    """voxelLods / aggregateLods false: no voxel level, no piece-by-piece group, no added vertex."""
    pos, uv, idx, _ = foliage.leaf_cards(fronds=4)
    d = dag.build_virtual_mesh(pos, idx, voxelLods=False, aggregateLods=False)
    assert d.voxelLevels == 0 and d.aggregateGroups == 0
    assert len(d.positions) == len(pos)
    check_dag("leaf cards, levels off", d, pos, idx)


def quad_row(n: int, spacing: float = 2.0):
    """n unit squares in the xz-plane along x, `spacing` apart, each its own piece (two triangles)."""
    pos, idx = [], []
    for i in range(n):
        x = spacing * i
        b = len(pos) // 3
        pos += [x - .5, 0, -.5, x + .5, 0, -.5, x + .5, 0, .5, x - .5, 0, .5]
        idx += [b, b + 2, b + 1, b, b + 3, b + 2]
    return np.array(pos, np.float32), np.array(idx, np.uint32)


def test_pair_thinning_closed_form():
    """aggregate.ts:155-205 on 1024 equal squares 2 apart: locked borders stall every group, each square is paired
    with its neighbour (Morton order along the row), the larger-or-equal one is kept, widened by sqrt 2 in its plane
    and moved to the pair's centre. The thinned error is the distance times max(w, 1 - w) = 2 x 1/2 = 1, above the
    widening (sqrt 2 - 1) / 2 and the bounding-box shrink 1 - sqrt 2 / 2 + 1/2; so level 1 holds half the triangles
    at error exactly 1, made of squares of side sqrt 2 centred at x = 2i + 1."""
    pos, idx = quad_row(1024)
    d = dag.build_virtual_mesh(pos, idx)
    assert d.aggregateGroups == 3, d.aggregateGroups  # 32 leaf meshlets in groups of 12, 12 and 8
    info = d.meshletInfo.reshape(-1, 4)
    b = d.meshletBounds.reshape(-1, 16)
    f0, c0, f1, c1 = (int(x) for x in d.levelRanges[:4])
    assert int(info[f1:f1 + c1, 0].sum()) == 1024
    assert np.all(b[f0:f0 + c0, 9] == 1.0) and np.all(b[f1:f1 + c1, 8] == 1.0)
    added = d.positions[len(pos):].reshape(-1, 3).astype(np.float64)
    assert len(added) == 2048
    h = math.sqrt(2) / 2
    centre = np.round((added[:, 0] - 1) / 2) * 2 + 1
    assert np.all(np.isin(np.round((centre - 1) / 2), np.arange(512) * 2))
    ulp = np.spacing(np.abs(added[:, 0]).astype(np.float32)).astype(np.float64)  # positions are binary32
    assert np.all(np.abs(np.abs(added[:, 0] - centre) - h) <= ulp + 1e-7)
    assert np.allclose(np.abs(added[:, 2]), h, atol=1e-6) and np.all(added[:, 1] == 0)
    check_dag("quad row", d, pos, idx)


def _edge_uses(tris):
    t = tris.reshape(-1, 3)
    e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), 1)
    return collections.Counter(map(tuple, e.tolist()))


def test_voxel_proxy_uvs_and_closed_surface():
    """import.test.ts:63-70 (every proxy vertex gets a finite source UV in [0, 1]); the surface-nets surface is
    closed (every edge used an even number of times), consistently oriented (an edge is used once in each
    direction per pair of faces), and within VOXEL_ERROR_CELLS cells of the unit sphere it was built from."""
    pos, uv, idx = foliage.sphere_geometry(1, 24, 12)
    pr = dag.build_voxel_proxy(pos, None, idx, 16, uvs=uv)
    u = pr["uvs"]
    assert len(u) == len(pr["positions"]) // 3 * 2
    assert np.all(np.isfinite(u)) and u.min() >= -1e-6 and u.max() <= 1 + 1e-6
    assert pr["colors"] is None
    assert all(c % 2 == 0 for c in _edge_uses(pr["indices"]).values()), "closed"
    t = pr["indices"].reshape(-1, 3)
    directed = collections.Counter(map(tuple, np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]).tolist()))
    assert all(directed[(a, b)] == directed[(b, a)] for (a, b) in directed), "consistently oriented"
    r = np.linalg.norm(pr["positions"].reshape(-1, 3), axis=1)
    assert pr["cellSize"] == 2.0 / 16
    assert np.max(np.abs(r - 1)) <= 2 * pr["cellSize"]


def test_voxel_proxy_colors_and_empty_inputs():
    """Proxy colours are means of the sampled source colours (a uniform colour stays that colour); no triangles
    or a zero extent give no proxy (voxelProxy.ts:40, 52)."""
    pos, _, idx = foliage.sphere_geometry(1, 16, 8)
    col = np.tile(np.array([0.25, 0.5, 0.75], np.float32), len(pos) // 3)
    pr = dag.build_voxel_proxy(pos, col, idx, 8)
    assert np.allclose(pr["colors"].reshape(-1, 3), [0.25, 0.5, 0.75], atol=1e-6)
    assert dag.build_voxel_proxy(pos, None, np.zeros(0, np.uint32), 8) is None
    assert dag.build_voxel_proxy(np.zeros(9, np.float32), None, np.array([0, 1, 2], np.uint32), 8) is None


def test_torus_uvs_cover_every_vertex():
    """import.test.ts:53-57 on TorusGeometry(2, 0.6, 48, 128), welded by position: leaf triangles kept, uvs for
    every vertex including voxel ones, roots under 1/50 of the leaf triangles."""
    pos, uv, idx = foliage.torus_geometry()
    wpos, widx, keep = foliage.weld_by_position(pos, idx)
    wuv = uv.reshape(-1, 2)[keep].reshape(-1)
    d = dag.build_virtual_mesh(wpos, widx, uvs=wuv)
    assert d.leafTriangles == len(idx) // 3
    assert d.uvs is not None and len(d.uvs) == len(d.positions) // 3 * 2
    assert d.rootTriangles < d.leafTriangles / 50, d.rootTriangles
    check_dag("torus", d, wpos, widx)


def test_conifer_like_crown_has_voxel_levels():
    """serialize.test.ts:68-71: a conifer's stacked separate cones get voxel LOD levels; the DAG stays valid and
    cuts stay complete."""
    pos, idx = foliage.stacked_cones()
    d = dag.build_virtual_mesh(pos, idx)
    assert d.voxelLevels > 0
    check_dag("conifer-like", d, pos, idx)
    for distance in (10, 100, 1000):
        vm = np.eye(4)
        vm[3, 1], vm[3, 2] = -8, -distance
        sel = cut64.select_cut(d, {"viewMatrix": vm.reshape(-1), "projScale": 1000.0, "near": 0.1}, 1.0)
        assert not cut64.verify_cut_coverage(d, sel) and not cut64.verify_cut_overlap(d, sel)


def test_replayed_runs_equal_fresh_runs(tmp_path):
    """The builder simplifies each voxel proxy once and replays that run to every target the levels ask for;
    tests/dag_replay_check.cpp compares the replay with a fresh run (indices and error) over three sources and
    their proxies at ~100 targets each, and a replay one collapse short must differ."""
    exe = str(tmp_path / "replay_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-subobject-linkage", "-pthread", "-o", exe,
                    os.path.join(ROOT, "tests", "dag_replay_check.cpp")], check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["cases"] > 200 and out["mismatches"] == 0, out
    assert out["planted_mismatches"] > 0.9 * out["cases"], out
