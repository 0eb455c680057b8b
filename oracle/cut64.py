"""ORACLE — test infrastructure only. fp64 restatement of the reference's CPU cut
(src/core/runtime/cut.ts:26-215) and lodFadeRange (src/core/runtime/lodBlend.ts:52-61).

These are the reference's own CPU statements of the LOD selection (no culling, one instance,
``view.viewMatrix`` = view * model, ``scale`` = uniform instance scale). The fp32 oracle (vgoracle.c)
is pinned against them: both must agree on every cluster not within 1e-6 relative of a threshold.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

ERROR_INFINITY = 1e30  # constants.ts:8
FAST_PATH_MAX_MESHLETS = 4  # constants.ts:14
FAST_PATH_MARGIN = 1e-3  # constants.ts:20
LOD_FADE_STEPS = 255  # lodBlend.ts:22


def _projected(data, view: dict, own: bool) -> np.ndarray:
    """projectedError of cut.ts:32-39 for every meshlet (own or parent sphere)."""
    v = np.asarray(view["viewMatrix"], dtype=np.float64)
    B = np.asarray(data.meshletBounds, dtype=np.float64).reshape(-1, 16)
    o = 0 if own else 4
    cx, cy, cz, r = B[:, o], B[:, o + 1], B[:, o + 2], B[:, o + 3]
    e = B[:, 8] if own else B[:, 9]
    s = view.get("scale", 1.0)
    x = v[0] * cx + v[4] * cy + v[8] * cz + v[12]
    y = v[1] * cx + v[5] * cy + v[9] * cz + v[13]
    z = v[2] * cx + v[6] * cy + v[10] * cz + v[14]
    d = np.maximum(np.sqrt(x * x + y * y + z * z) - r * s, view["near"])
    with np.errstate(over="ignore"):
        px = (e * s * view["projScale"]) / d
    return np.where(e >= ERROR_INFINITY * 0.5, np.inf, px)


def select_cut(data, view: dict, threshold: float) -> np.ndarray:
    """selectCut, cut.ts:26-48: own <= t < parent."""
    own = _projected(data, view, True)
    parent = _projected(data, view, False)
    return ((own <= threshold) & (parent > threshold)).astype(np.uint8)


def lod_fade_range(own: float, parent: float, threshold: float, blend: float) -> Optional[Tuple[int, int]]:
    """lodFadeRange, lodBlend.ts:52-61 (Math.round: half toward +infinity)."""
    high = threshold * blend if blend > 1 else threshold
    if not (own <= high and parent > threshold):
        return None
    if not (blend > 1):
        return (0, LOD_FADE_STEPS)
    scale = LOD_FADE_STEPS / math.log(blend)

    def step(e):
        return min(LOD_FADE_STEPS, max(0, math.floor(math.log(max(e, 1e-30) / threshold) * scale + 0.5)))

    lo = step(own)
    hi = step(parent) if math.isfinite(parent) else LOD_FADE_STEPS
    return (lo, hi) if hi > lo else None


def select_blend_cut(data, view: dict, threshold: float, blend: float):
    """selectBlendCut, cut.ts:56-77: per meshlet the blend steps [lo, hi) it draws (hi = 0: not drawn)."""
    own = _projected(data, view, True)
    parent = _projected(data, view, False)
    lo = np.zeros(data.meshletCount, dtype=np.uint8)
    hi = np.zeros(data.meshletCount, dtype=np.uint8)
    for i in range(data.meshletCount):
        r = lod_fade_range(float(own[i]), float(parent[i]), threshold, blend)
        if r:
            lo[i], hi[i] = r
    return lo, hi


def verify_cut_overlap(data, selected: np.ndarray) -> list:
    """verifyCutOverlap, cut.ts:83-97: selected meshlets with a selected coarser replacement (recursively)."""
    group, start, lst = data.replacementGroup, data.replacementStart, data.replacementIndices
    info = np.asarray(data.meshletInfo).reshape(-1, 4)
    coarser = np.zeros(data.meshletCount, dtype=bool)
    # Replacements are on coarser (higher) levels: resolve from the coarsest level down.
    for i in np.argsort(-info[:, 2].astype(np.int64), kind="stable"):
        g = group[i]
        if g >= 0:
            ks = lst[start[g]:start[g + 1]]
            coarser[i] = bool(np.any(selected[ks] == 1) or np.any(coarser[ks]))
    return [int(i) for i in np.nonzero((selected == 1) & coarser)[0]]


def verify_cut_coverage(data, selected: np.ndarray) -> list:
    """verifyCutCoverage, cut.ts:104-122: leaves not covered by a selected meshlet on their path up."""
    group, start, lst = data.replacementGroup, data.replacementStart, data.replacementIndices
    info = np.asarray(data.meshletInfo).reshape(-1, 4)
    covered = np.full(data.meshletCount, -1, dtype=np.int8)
    # Replacements are coarser (higher level), so resolve from the coarsest level down.
    levels = info[:, 2]
    for i in np.argsort(-levels.astype(np.int64), kind="stable"):
        res = selected[i] == 1
        g = group[i]
        if not res and g >= 0:
            res = any(covered[k] == 1 for k in lst[start[g]:start[g + 1]])
        covered[i] = 1 if res else 0
    return [i for i in range(data.meshletCount) if info[i, 2] == 0 and covered[i] != 1]


def lod_distance_bounds(data, view: dict) -> Tuple[float, float]:
    """cut.ts:125-134."""
    v = view["viewMatrix"]
    cx, cy, cz, r = (float(a) for a in data.lodBoundsSphere)
    x = v[0] * cx + v[4] * cy + v[8] * cz + v[12]
    y = v[1] * cx + v[5] * cy + v[9] * cz + v[13]
    z = v[2] * cx + v[6] * cy + v[10] * cz + v[14]
    d = math.sqrt(x * x + y * y + z * z)
    R = r * view.get("scale", 1.0)
    return max(d - R, view["near"]), max(d + R, view["near"])


def instance_meshlet_range(data, view: dict, threshold: float, blend: float = 1.0) -> Tuple[int, int]:
    """instanceMeshletRange, cut.ts:146-166."""
    d_min, d_max = lod_distance_bounds(data, view)
    high = threshold * max(1.0, blend)
    k = view.get("scale", 1.0) * view["projScale"]
    lo = hi = -1
    levels = len(data.levelRanges) // 2
    for lv in range(levels):
        min_own = float(data.levelErrors[lv * 2])
        max_parent = float(data.levelErrors[lv * 2 + 1])
        own_possible = (min_own * k) / d_max <= high
        parent_possible = max_parent >= ERROR_INFINITY * 0.5 or (max_parent * k) / d_min > threshold
        if own_possible and parent_possible:
            if lo < 0:
                lo = lv
            hi = lv
    if lo < 0:
        return (0, 0)
    return int(data.levelRanges[lo * 2]), int(data.levelRanges[hi * 2] + data.levelRanges[hi * 2 + 1])


def instance_guaranteed_range(data, view: dict, threshold: float, blend: float = 1.0) -> Optional[Tuple[int, int]]:
    """instanceGuaranteedRange, cut.ts:180-215."""
    d_min, d_max = lod_distance_bounds(data, view)
    high = threshold * max(1.0, blend)
    margin = FAST_PATH_MARGIN
    k = view.get("scale", 1.0) * view["projScale"]
    lo = hi = -1
    possible = guaranteed = 0
    levels = len(data.levelRanges) // 2
    for lv in range(levels):
        min_own = float(data.levelErrors[lv * 2])
        max_parent = float(data.levelErrors[lv * 2 + 1])
        own_possible = (min_own * k) / d_max <= high
        parent_possible = max_parent >= ERROR_INFINITY * 0.5 or (max_parent * k) / d_min > threshold
        if not own_possible or not parent_possible:
            continue
        if lo < 0:
            lo = lv
        hi = lv
        size = int(data.levelRanges[lv * 2 + 1])
        possible += size
        max_own = float(data.levelGuardErrors[lv * 2])
        min_parent = float(data.levelGuardErrors[lv * 2 + 1])
        own_sure = (max_own * k) / d_min <= threshold * (1 - margin)
        parent_sure = min_parent >= ERROR_INFINITY * 0.5 or (min_parent * k) / d_max > high * (1 + margin)
        if own_sure and parent_sure:
            guaranteed += size
    if lo < 0:
        return None
    start = int(data.levelRanges[lo * 2])
    end = int(data.levelRanges[hi * 2] + data.levelRanges[hi * 2 + 1])
    if possible != end - start or guaranteed != possible:
        return None
    if possible > FAST_PATH_MAX_MESHLETS:
        return None
    return (start, end)


def park_miller(seed: int):
    """The reference tests' seeded generator: seed * 16807 mod 2^31 - 1 (tests/cut.test.ts:11-12)."""
    state = [seed]

    def rand():
        state[0] = (state[0] * 16807) % 2147483647
        return state[0] / 2147483647

    return rand
