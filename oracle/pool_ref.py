"""ORACLE — test infrastructure only. Independent numpy restatement of the pool's host-side layout:
mortonOrder / radixOrder (GeometryPool.ts:1249-1311), instanceSphere (GeometryPool.ts:1171-1180),
computeCell (GeometryPool.ts:519-545), the mesh record and level table (GeometryPool.ts:432-460).
Column lengths are sqrt of the sum of squares (the library's choice for Math.hypot).
"""
from __future__ import annotations

import numpy as np

CELL_SIZE = 128


def instance_spheres(matrices: np.ndarray, sphere) -> np.ndarray:
    m = np.asarray(matrices, dtype=np.float32).reshape(-1, 16).astype(np.float64)
    sx, sy, sz, sr = (float(v) for v in sphere)
    out = np.empty((len(m), 4))
    out[:, 0] = m[:, 0] * sx + m[:, 4] * sy + m[:, 8] * sz + m[:, 12]
    out[:, 1] = m[:, 1] * sx + m[:, 5] * sy + m[:, 9] * sz + m[:, 13]
    out[:, 2] = m[:, 2] * sx + m[:, 6] * sy + m[:, 10] * sz + m[:, 14]
    cols = [np.sqrt(m[:, k * 4] * m[:, k * 4] + m[:, k * 4 + 1] * m[:, k * 4 + 1] + m[:, k * 4 + 2] * m[:, k * 4 + 2]) for k in range(3)]
    out[:, 3] = sr * np.maximum(cols[0], np.maximum(cols[1], cols[2]))
    return out


def _spread(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint32) & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def morton_order(matrices: np.ndarray, sphere) -> np.ndarray:
    """Caller index of each GPU slot: 30-bit Morton code, stable sort by code (index tie-break)."""
    s = instance_spheres(matrices, sphere)
    c = s[:, :3]
    lo, hi = c.min(axis=0), c.max(axis=0)
    extent = max(float(np.max(hi - lo)), 1e-9)
    q = np.clip(np.floor(((c - lo) / extent) * 1024), 0, 1023).astype(np.uint32)
    codes = _spread(q[:, 0]) | (_spread(q[:, 1]) << 1) | (_spread(q[:, 2]) << 2)
    return np.argsort(codes, kind="stable").astype(np.uint32)


def compute_cells(gpu_matrices: np.ndarray, sphere, slot: int, instance_base: int) -> np.ndarray:
    """8 floats per cell of CELL_SIZE consecutive GPU-ordered instances."""
    s = instance_spheres(gpu_matrices, sphere)
    n = len(s)
    cells = []
    for first in range(0, n, CELL_SIZE):
        end = min(first + CELL_SIZE, n)
        t = s[first:end]
        lo = (t[:, :3] - t[:, 3:4]).min(axis=0)
        hi = (t[:, :3] + t[:, 3:4]).max(axis=0)
        center = (lo + hi) / 2
        d = t[:, :3] - center
        radius = float(np.max(np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) + t[:, 3]))
        max_radius = float(t[:, 3].max())
        cells.append([center[0], center[1], center[2], radius * (1 + 1e-4) + 1e-3, max_radius * (1 + 1e-4) + 1e-4, slot,
                      instance_base + first, end - first])
    return np.asarray(cells, dtype=np.float32).reshape(-1)


def mesh_table(meshes, slots: int, flags=None) -> np.ndarray:
    """Mesh records (6 vec4 per slot) then each mesh's level table (2 vec4 per level), in add order."""
    table = np.zeros((slots * 6, 4), dtype=np.float32)
    levels = []
    level_used = 0
    meshlet_base = 0
    for slot, d in enumerate(meshes):
        L = len(d.levelRanges) // 2
        base = slots * 6 + level_used
        for lv in range(L):
            levels.append([d.levelRanges[lv * 2] + meshlet_base, d.levelRanges[lv * 2 + 1], d.levelErrors[lv * 2], d.levelErrors[lv * 2 + 1]])
            levels.append([d.levelGuardErrors[lv * 2], d.levelGuardErrors[lv * 2 + 1], 0, 0])
        rec = table[slot * 6:(slot + 1) * 6]
        rec[0] = d.boundingSphere
        rec[1] = d.lodBoundsSphere
        rec[2] = d.cullBoundsSphere
        f = (flags or {}).get(slot, (1e30, 0.0, 1, 1, 0))
        rec[3] = (f[0], f[1], L, base)
        rec[4] = (1, 1, 0, 0)
        rec[5] = (f[2], f[3], f[4], 0)
        level_used += 2 * L
        meshlet_base += d.meshletCount
    out = np.concatenate([table, np.asarray(levels, dtype=np.float32).reshape(-1, 4)]) if levels else table
    return out.reshape(-1)
