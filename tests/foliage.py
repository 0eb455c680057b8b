"""Test meshes made of separate pieces, restating the reference's test assets.

leaf_cards: tests/leafCards.ts (a palm-like crown of bent three-quad leaflets), with three.js's Vector3 arithmetic
in double precision and positions stored as binary32 as Float32BufferAttribute does.
"""
from __future__ import annotations

import math

import numpy as np


def _vertex_normals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """BufferGeometry.computeVertexNormals: per-vertex sum of face normals (cb x ab), normalised."""
    p = pos.reshape(-1, 3).astype(np.float64)
    t = idx.reshape(-1, 3)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    fn = np.cross(c - b, a - b)
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, t[:, k], fn)
    ln = np.linalg.norm(n, axis=1)
    ln[ln == 0] = 1
    return (n / ln[:, None]).astype(np.float32).reshape(-1)


def leaf_cards(fronds: int = 12, leaflets_per_side: int = 36, seed: int = 3):
    """tests/leafCards.ts leafCards(): positions, uvs, indices (uint32) and normals."""
    state = [seed]

    def rand():
        state[0] = (state[0] * 16807) % 2147483647
        return state[0] / 2147483647

    positions, uvs, indices = [], [], []
    for f in range(fronds):
        yaw = (f / fronds) * math.pi * 2 + rand() * 0.3
        d = (math.cos(yaw), 0.0, math.sin(yaw))
        side = (-d[2], 0.0, d[0])
        for i in range(leaflets_per_side):
            t = (i + 0.5) / leaflets_per_side
            s = t * 4.5
            base = (d[0] * s + 0, d[1] * s + (8 + 1.2 * t - 2.2 * t * t), d[2] * s + 0)
            for sign in (-1, 1):
                o = [side[k] * sign + d[k] * 0.5 for k in range(3)]
                ln = math.sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) or 1
                inv = 1 / ln
                o = [x * inv for x in o]
                length = 1.1 * (1 - 0.6 * abs(t - 0.4)) * (0.85 + 0.3 * rand())
                first = len(positions) // 3
                for k in range(4):
                    u = k / 3
                    p = [base[j] + o[j] * (u * length) for j in range(3)]
                    p[1] -= 0.25 * u * u * length
                    for w in (-0.033, 0.033):
                        positions += [p[0] + d[0] * w, p[1] + d[1] * w, p[2] + d[2] * w]
                        uvs += [u, 1 if w > 0 else 0]
                for k in range(3):
                    a = first + k * 2
                    indices += [a, a + 1, a + 2, a + 1, a + 3, a + 2]
    pos = np.array(positions, dtype=np.float32)
    idx = np.array(indices, dtype=np.uint32)
    return pos, np.array(uvs, dtype=np.float32), idx, _vertex_normals(pos, idx)


def sphere_geometry(radius: float = 1.0, width_segments: int = 32, height_segments: int = 16):
    """THREE.SphereGeometry's vertex grid (seam and pole copies kept): positions, uvs, indices."""
    pos, uv, grid = [], [], []
    for iy in range(height_segments + 1):
        v = iy / height_segments
        u_off = 0.5 / width_segments if iy == 0 else (-0.5 / width_segments if iy == height_segments else 0)
        row = []
        for ix in range(width_segments + 1):
            u = ix / width_segments
            pos += [-radius * math.cos(u * 2 * math.pi) * math.sin(v * math.pi), radius * math.cos(v * math.pi),
                    radius * math.sin(u * 2 * math.pi) * math.sin(v * math.pi)]
            uv += [u + u_off, 1 - v]
            row.append(len(pos) // 3 - 1)
        grid.append(row)
    idx = []
    for iy in range(height_segments):
        for ix in range(width_segments):
            a, b, c, d = grid[iy][ix + 1], grid[iy][ix], grid[iy + 1][ix], grid[iy + 1][ix + 1]
            if iy != 0:
                idx += [a, b, d]
            if iy != height_segments - 1:
                idx += [b, c, d]
    return np.array(pos, np.float32), np.array(uv, np.float32), np.array(idx, np.uint32)


def torus_geometry(radius: float = 2.0, tube: float = 0.6, radial_segments: int = 48, tubular_segments: int = 128):
    """THREE.TorusGeometry's vertex grid (seam copies kept): positions, uvs, indices."""
    pos, uv = [], []
    for j in range(radial_segments + 1):
        for i in range(tubular_segments + 1):
            u = i / tubular_segments * math.pi * 2
            v = j / radial_segments * math.pi * 2
            pos += [(radius + tube * math.cos(v)) * math.cos(u), (radius + tube * math.cos(v)) * math.sin(u),
                    tube * math.sin(v)]
            uv += [i / tubular_segments, j / radial_segments]
    idx = []
    for j in range(1, radial_segments + 1):
        for i in range(1, tubular_segments + 1):
            a = (tubular_segments + 1) * j + i - 1
            b = (tubular_segments + 1) * (j - 1) + i - 1
            c = (tubular_segments + 1) * (j - 1) + i
            d = (tubular_segments + 1) * j + i
            idx += [a, b, d, b, c, d]
    return np.array(pos, np.float32), np.array(uv, np.float32), np.array(idx, np.uint32)


def weld_by_position(pos: np.ndarray, idx: np.ndarray):
    """Seam copies merged (the first copy of each position kept), as a position-only weld."""
    p = pos.reshape(-1, 3)
    _, first, inverse = np.unique(p, axis=0, return_index=True, return_inverse=True)
    keep = np.sort(first)
    new_id = np.full(len(p), -1, np.int64)
    new_id[keep] = np.arange(len(keep))
    remap = new_id[first[inverse.reshape(-1)]]
    return p[keep].reshape(-1).astype(np.float32), remap[idx].astype(np.uint32), keep


def stacked_cones(tiers: int = 8, segments: int = 48, rings: int = 8, seed: int = 5):
    """A conifer-like crown: a closed trunk cylinder and `tiers` separate closed cones stacked with overlap and
    drooping, wavy rims (the shape of the reference demo's makeConifer, examples/demo/assets.ts:81-123, without its
    Perlin noise). Each part is its own closed piece, so simplification can only shrink the silhouette."""
    rng = np.random.default_rng(seed)
    parts_pos, parts_idx, base = [], [], 0

    def add(p, i):
        nonlocal base
        parts_pos.append(p)
        parts_idx.append(i + base)
        base += len(p)

    def lathe(profile, phase, wobble):
        """Closed surface of revolution through (radius, y) rings, capped by the first and last points."""
        n = len(profile)
        ang = np.linspace(0, 2 * np.pi, segments, endpoint=False)
        pts = []
        for r, y in profile:
            w = 1 + wobble * np.sin(ang * 7 + phase) * (r / max(profile[0][0], 1e-6))
            pts.append(np.stack([r * w * np.cos(ang), np.full(segments, y) - 0.7 * (r * w / max(profile[0][0], 1e-6)) ** 2 * wobble * 4,
                                 r * w * np.sin(ang)], 1))
        p = np.concatenate(pts + [np.array([[0, profile[0][1], 0], [0, profile[-1][1], 0]])], 0)
        idx = []
        for k in range(n - 1):
            for s in range(segments):
                a, b = k * segments + s, k * segments + (s + 1) % segments
                c, d = a + segments, b + segments
                idx += [a, c, b, b, c, d]
        bottom, top = n * segments, n * segments + 1
        for s in range(segments):
            idx += [bottom, s, (s + 1) % segments]
            t0 = (n - 1) * segments
            idx += [top, t0 + (s + 1) % segments, t0 + s]
        return p, np.array(idx, np.int64)

    height = 16.0
    trunk = [(0.42 - 0.24 * k / rings, height * 0.8 * k / rings) for k in range(rings + 1)]
    add(*lathe(trunk, 0.0, 0.0))
    for i in range(tiers):
        t = i / (tiers - 1)
        radius = 4.4 + (0.9 - 4.4) * t
        cone_h = 4.6 + (2.8 - 4.6) * t
        y0 = 2.6 + t * (height - 4.4)
        prof = [(radius * (1 - k / rings) + 1e-3, y0 + cone_h * k / rings) for k in range(rings + 1)]
        add(*lathe(prof, float(rng.uniform(0, 6.28)), 0.12))
    pos = np.concatenate(parts_pos, 0).astype(np.float32).reshape(-1)
    idx = np.concatenate(parts_idx, 0).astype(np.uint32)
    return pos, idx
