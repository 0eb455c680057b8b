"""Meshlet LOD DAG builder for indexed triangle meshes (lib/libvgdag.so, include/vgdag.h).

Python face of the reference's ``buildVirtualMesh(src, options)`` (src/core/preprocess/buildVirtualMesh.ts:185)
and ``partitionMeshlets(meshlets, groupSize)`` (src/core/preprocess/partition.ts:16), with the reference's
option names. The builder needs no meshoptimizer; its voxel-proxy and piece-by-piece aggregate levels
(options ``voxelLods``, ``aggregateLods``) are not built, so it behaves as the reference with both off.
"""
from __future__ import annotations

import ctypes
from typing import Iterable, List, Sequence

import numpy as np

from . import _native
from .scenegen import VirtualMeshData


class _Options(ctypes.Structure):
    _fields_ = [
        ("group_size", ctypes.c_uint32),
        ("max_lod_levels", ctypes.c_uint32),
        ("max_group_keep_ratio", ctypes.c_double),
        ("simplify_ratio", ctypes.c_double),
        ("min_level_reduction", ctypes.c_double),
        ("prune", ctypes.c_int32),
        ("tail_simplify", ctypes.c_int32),
        ("min_root_triangles", ctypes.c_uint32),
        ("threads", ctypes.c_uint32),
    ]


class _View(ctypes.Structure):
    _fields_ = [
        ("vertex_count", ctypes.c_uint32),
        ("positions", ctypes.POINTER(ctypes.c_float)),
        ("index_count", ctypes.c_uint64),
        ("indices", ctypes.POINTER(ctypes.c_uint32)),
        ("meshlet_count", ctypes.c_uint32),
        ("meshlet_bounds", ctypes.POINTER(ctypes.c_float)),
        ("meshlet_info", ctypes.POINTER(ctypes.c_uint32)),
        ("meshlet_vertex_count", ctypes.c_uint64),
        ("meshlet_vertices", ctypes.POINTER(ctypes.c_uint32)),
        ("meshlet_triangles", ctypes.POINTER(ctypes.c_uint32)),
        ("level_count", ctypes.c_uint32),
        ("level_ranges", ctypes.POINTER(ctypes.c_uint32)),
        ("level_errors", ctypes.POINTER(ctypes.c_float)),
        ("level_guard_errors", ctypes.POINTER(ctypes.c_float)),
        ("replacement_group", ctypes.POINTER(ctypes.c_int32)),
        ("group_count", ctypes.c_uint32),
        ("replacement_start", ctypes.POINTER(ctypes.c_uint32)),
        ("replacement_indices", ctypes.POINTER(ctypes.c_uint32)),
        ("bounding_sphere", ctypes.c_float * 4),
        ("lod_bounds_sphere", ctypes.c_float * 4),
        ("cull_bounds_sphere", ctypes.c_float * 4),
        ("leaf_triangles", ctypes.c_uint64),
        ("leaf_meshlets", ctypes.c_uint32),
        ("root_triangles", ctypes.c_uint32),
        ("root_meshlets", ctypes.c_uint32),
        ("lod_levels", ctypes.c_uint32),
    ]


_lib = None
EXPORTS = ("vgd_default_options", "vgd_build", "vgd_mesh_get", "vgd_mesh_free", "vgd_partition")


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(_native.library_path("libvgdag.so"))
        lib.vgd_default_options.argtypes = [ctypes.POINTER(_Options)]
        lib.vgd_default_options.restype = None
        lib.vgd_build.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(_Options),
                                  ctypes.POINTER(ctypes.c_void_p)]
        lib.vgd_build.restype = ctypes.c_int
        lib.vgd_mesh_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(_View)]
        lib.vgd_mesh_get.restype = None
        lib.vgd_mesh_free.argtypes = [ctypes.c_void_p]
        lib.vgd_mesh_free.restype = None
        lib.vgd_partition.argtypes = [ctypes.c_uint32] + [ctypes.c_void_p] * 4 + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
        lib.vgd_partition.restype = ctypes.c_uint32
        _lib = lib
    return _lib


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


_OPTION_NAMES = {
    "groupSize": "group_size",
    "maxLodLevels": "max_lod_levels",
    "maxGroupKeepRatio": "max_group_keep_ratio",
    "simplifyRatio": "simplify_ratio",
    "minLevelReduction": "min_level_reduction",
    "prune": "prune",
    "tailSimplify": "tail_simplify",
    "minRootTriangles": "min_root_triangles",
    "threads": "threads",
}
_UNBUILT = ("voxelLods", "aggregateLods", "voxelResolution")


def build_virtual_mesh(positions, indices, **options) -> VirtualMeshData:
    """buildVirtualMesh (buildVirtualMesh.ts:185) over float32 xyz positions and uint32 triangle indices.

    Options take the reference's names (``groupSize``, ``maxLodLevels``, ``maxGroupKeepRatio``,
    ``simplifyRatio``, ``minLevelReduction``, ``prune``, ``tailSimplify``, ``minRootTriangles``) plus
    ``threads``. ``voxelLods`` / ``aggregateLods`` are accepted only as False.
    """
    lib = _load()
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(-1))
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
    opt = _Options()
    lib.vgd_default_options(ctypes.byref(opt))
    for k, v in options.items():
        if k in _UNBUILT:
            if v:
                raise ValueError(f"{k} is not supported by this builder")
            continue
        if k not in _OPTION_NAMES:
            raise TypeError(f"unknown option {k}")
        field = _OPTION_NAMES[k]
        setattr(opt, field, int(v) if field in ("prune", "tail_simplify") else v)
    handle = ctypes.c_void_p()
    rc = lib.vgd_build(pos.ctypes.data, len(pos) // 3, idx.ctypes.data, len(idx), ctypes.byref(opt), ctypes.byref(handle))
    if rc != 0:
        raise ValueError(f"vgd_build failed ({rc})")
    try:
        v = _View()
        lib.vgd_mesh_get(handle, ctypes.byref(v))
        n, levels, groups = v.meshlet_count, v.level_count, v.group_count
        start = _arr(v.replacement_start, groups + 1, np.uint32)
        return VirtualMeshData(
            meshletCount=n,
            meshletBounds=_arr(v.meshlet_bounds, n * 16, np.float32),
            meshletInfo=_arr(v.meshlet_info, n * 4, np.uint32),
            levelRanges=_arr(v.level_ranges, levels * 2, np.uint32),
            levelErrors=_arr(v.level_errors, levels * 2, np.float32),
            levelGuardErrors=_arr(v.level_guard_errors, levels * 2, np.float32),
            replacementGroup=_arr(v.replacement_group, n, np.int32),
            replacementStart=start,
            replacementIndices=_arr(v.replacement_indices, int(start[-1]), np.uint32),
            boundingSphere=np.array(v.bounding_sphere, dtype=np.float32),
            lodBoundsSphere=np.array(v.lod_bounds_sphere, dtype=np.float32),
            cullBoundsSphere=np.array(v.cull_bounds_sphere, dtype=np.float32),
            leafTriangles=int(v.leaf_triangles),
            leafMeshlets=int(v.leaf_meshlets),
            rootTriangles=int(v.root_triangles),
            rootMeshlets=int(v.root_meshlets),
            meshletTriangles=_arr(v.meshlet_triangles, int(v.index_count) // 3, np.uint32),
            positions=_arr(v.positions, v.vertex_count * 3, np.float32),
            indices=_arr(v.indices, int(v.index_count), np.uint32),
            meshletVertices=_arr(v.meshlet_vertices, int(v.meshlet_vertex_count), np.uint32),
            lodLevels=int(v.lod_levels),
        )
    finally:
        lib.vgd_mesh_free(handle)


def edge_key(a: int, b: int) -> int:
    """Order-independent edge id of two welded vertex ids (buildVirtualMesh.ts:684)."""
    return a * 67108864 + b if a < b else b * 67108864 + a


def partition_meshlets(meshlets: Sequence[dict], group_size: int) -> List[List[int]]:
    """partitionMeshlets (partition.ts:16): ``meshlets`` hold ``boundaryEdges`` (edge keys) and ``center``."""
    lib = _load()
    n = len(meshlets)
    counts = np.array([len(m["boundaryEdges"]) for m in meshlets], dtype=np.uint32)
    offsets = np.zeros(n, dtype=np.uint32)
    if n:
        offsets[1:] = np.cumsum(counts)[:-1]
    edges = np.array([e for m in meshlets for e in m["boundaryEdges"]] or [0], dtype=np.uint64)
    centers = np.ascontiguousarray(np.array([m["center"] for m in meshlets] or [[0, 0, 0]], dtype=np.float64).reshape(-1))
    group_of = np.zeros(max(n, 1), dtype=np.uint32)
    members = np.zeros(max(n, 1), dtype=np.uint32)
    groups = lib.vgd_partition(n, edges.ctypes.data, offsets.ctypes.data, counts.ctypes.data, centers.ctypes.data, group_size,
                               group_of.ctypes.data, members.ctypes.data)
    out: List[List[int]] = [[] for _ in range(groups)]
    for i in members[:n]:
        out[int(group_of[i])].append(int(i))
    return out
