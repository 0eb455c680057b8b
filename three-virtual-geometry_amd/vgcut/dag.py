"""Meshlet LOD DAG builder for indexed triangle meshes (lib/libvgdag.so, include/vgdag.h).

Python face of the reference's ``buildVirtualMesh(src, options)`` (src/core/preprocess/buildVirtualMesh.ts:185)
and ``partitionMeshlets(meshlets, groupSize)`` (src/core/preprocess/partition.ts:16), with the reference's
option names and defaults, including the voxel-proxy tail levels (``voxelLods``, ``voxelResolution``;
voxelProxy.ts) and the piece-by-piece reduction of groups made of separate pieces (``aggregateLods``;
aggregate.ts). The builder needs no meshoptimizer.
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
        ("voxel_lods", ctypes.c_int32),
        ("voxel_resolution", ctypes.c_uint32),
        ("aggregate_lods", ctypes.c_int32),
    ]


class _Source(ctypes.Structure):
    _fields_ = [
        ("vertex_count", ctypes.c_uint32),
        ("positions", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("colors", ctypes.c_void_p),
        ("uvs", ctypes.c_void_p),
        ("index_count", ctypes.c_uint64),
        ("indices", ctypes.c_void_p),
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
        ("normals", ctypes.POINTER(ctypes.c_float)),
        ("colors", ctypes.POINTER(ctypes.c_float)),
        ("uvs", ctypes.POINTER(ctypes.c_float)),
        ("voxel_levels", ctypes.c_uint32),
        ("aggregate_groups", ctypes.c_uint32),
    ]


class _ProxyView(ctypes.Structure):
    _fields_ = [
        ("vertex_count", ctypes.c_uint32),
        ("positions", ctypes.POINTER(ctypes.c_float)),
        ("normals", ctypes.POINTER(ctypes.c_float)),
        ("colors", ctypes.POINTER(ctypes.c_float)),
        ("uvs", ctypes.POINTER(ctypes.c_float)),
        ("index_count", ctypes.c_uint64),
        ("indices", ctypes.POINTER(ctypes.c_uint32)),
        ("cell_size", ctypes.c_double),
    ]


_lib = None
EXPORTS = ("vgd_default_options", "vgd_build", "vgd_build_source", "vgd_mesh_get", "vgd_mesh_free", "vgd_partition",
           "vgd_voxel_proxy", "vgd_proxy_get", "vgd_proxy_free", "vgd_vertex_encode_bound", "vgd_vertex_encode",
           "vgd_vertex_decode", "vgd_index_encode_bound", "vgd_index_encode", "vgd_index_decode")


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(_native.library_path("libvgdag.so"))
        lib.vgd_default_options.argtypes = [ctypes.POINTER(_Options)]
        lib.vgd_default_options.restype = None
        lib.vgd_build.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(_Options),
                                  ctypes.POINTER(ctypes.c_void_p)]
        lib.vgd_build.restype = ctypes.c_int
        lib.vgd_build_source.argtypes = [ctypes.POINTER(_Source), ctypes.POINTER(_Options), ctypes.POINTER(ctypes.c_void_p)]
        lib.vgd_build_source.restype = ctypes.c_int
        lib.vgd_mesh_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(_View)]
        lib.vgd_mesh_get.restype = None
        lib.vgd_mesh_free.argtypes = [ctypes.c_void_p]
        lib.vgd_mesh_free.restype = None
        lib.vgd_partition.argtypes = [ctypes.c_uint32] + [ctypes.c_void_p] * 4 + [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
        lib.vgd_partition.restype = ctypes.c_uint32
        lib.vgd_voxel_proxy.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        lib.vgd_voxel_proxy.restype = ctypes.c_int
        lib.vgd_proxy_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(_ProxyView)]
        lib.vgd_proxy_get.restype = None
        lib.vgd_proxy_free.argtypes = [ctypes.c_void_p]
        lib.vgd_proxy_free.restype = None
        sz, vp = ctypes.c_size_t, ctypes.c_void_p
        lib.vgd_vertex_encode_bound.argtypes = [sz, sz]
        lib.vgd_vertex_encode_bound.restype = sz
        lib.vgd_vertex_encode.argtypes = [vp, sz, vp, sz, sz]
        lib.vgd_vertex_encode.restype = sz
        lib.vgd_vertex_decode.argtypes = [vp, sz, sz, vp, sz]
        lib.vgd_vertex_decode.restype = ctypes.c_int
        lib.vgd_index_encode_bound.argtypes = [sz, sz]
        lib.vgd_index_encode_bound.restype = sz
        lib.vgd_index_encode.argtypes = [vp, sz, vp, sz]
        lib.vgd_index_encode.restype = sz
        lib.vgd_index_decode.argtypes = [vp, sz, vp, sz]
        lib.vgd_index_decode.restype = ctypes.c_int
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
    "voxelLods": "voxel_lods",
    "voxelResolution": "voxel_resolution",
    "aggregateLods": "aggregate_lods",
}
_FLAGS = ("prune", "tail_simplify", "voxel_lods", "aggregate_lods")


def _attribute(values, vertices: int, size: int, name: str):
    if values is None:
        return None
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float32).reshape(-1))
    if len(a) != vertices * size:
        raise ValueError(f"{name} must hold {size} floats per vertex")
    return a


def build_virtual_mesh(positions, indices, normals=None, colors=None, uvs=None, **options) -> VirtualMeshData:
    """buildVirtualMesh (buildVirtualMesh.ts:185) over float32 xyz positions and uint32 triangle indices.

    ``normals`` / ``colors`` (3 floats per vertex) and ``uvs`` (2) are optional; when given, the result
    carries them for every vertex, including those that thinned pieces and voxel levels add. Options take
    the reference's names (``groupSize``, ``maxLodLevels``, ``maxGroupKeepRatio``, ``simplifyRatio``,
    ``minLevelReduction``, ``prune``, ``tailSimplify``, ``minRootTriangles``, ``voxelLods``,
    ``voxelResolution``, ``aggregateLods``) plus ``threads``.
    """
    lib = _load()
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(-1))
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
    nv = len(pos) // 3
    nrm = _attribute(normals, nv, 3, "normals")
    col = _attribute(colors, nv, 3, "colors")
    uv = _attribute(uvs, nv, 2, "uvs")
    opt = _Options()
    lib.vgd_default_options(ctypes.byref(opt))
    for k, v in options.items():
        if k not in _OPTION_NAMES:
            raise TypeError(f"unknown option {k}")
        field = _OPTION_NAMES[k]
        setattr(opt, field, int(bool(v)) if field in _FLAGS else v)
    src = _Source(nv, pos.ctypes.data, nrm.ctypes.data if nrm is not None else None,
                  col.ctypes.data if col is not None else None, uv.ctypes.data if uv is not None else None,
                  len(idx), idx.ctypes.data)
    handle = ctypes.c_void_p()
    rc = lib.vgd_build_source(ctypes.byref(src), ctypes.byref(opt), ctypes.byref(handle))
    if rc != 0:
        raise ValueError(f"vgd_build failed ({rc})")
    try:
        v = _View()
        lib.vgd_mesh_get(handle, ctypes.byref(v))
        n, levels, groups = v.meshlet_count, v.level_count, v.group_count
        nv_out = v.vertex_count
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
            normals=_arr(v.normals, nv_out * 3, np.float32) if v.normals else None,
            colors=_arr(v.colors, nv_out * 3, np.float32) if v.colors else None,
            uvs=_arr(v.uvs, nv_out * 2, np.float32) if v.uvs else None,
            voxelLevels=int(v.voxel_levels),
            aggregateGroups=int(v.aggregate_groups),
        )
    finally:
        lib.vgd_mesh_free(handle)


def build_voxel_proxy(positions, colors, indices, resolution: int, uvs=None):
    """buildVoxelProxy (voxelProxy.ts:33): a dict of positions, normals, colors, uvs, indices and cellSize, or None
    when there is no proxy."""
    lib = _load()
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(-1))
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
    nv = len(pos) // 3
    col = _attribute(colors, nv, 3, "colors")
    uv = _attribute(uvs, nv, 2, "uvs")
    handle = ctypes.c_void_p()
    rc = lib.vgd_voxel_proxy(pos.ctypes.data, nv, col.ctypes.data if col is not None else None,
                             uv.ctypes.data if uv is not None else None, idx.ctypes.data, len(idx), int(resolution),
                             ctypes.byref(handle))
    if rc == 1:
        return None
    if rc != 0:
        raise ValueError(f"vgd_voxel_proxy failed ({rc})")
    try:
        v = _ProxyView()
        lib.vgd_proxy_get(handle, ctypes.byref(v))
        n = v.vertex_count
        return {
            "positions": _arr(v.positions, n * 3, np.float32),
            "normals": _arr(v.normals, n * 3, np.float32),
            "colors": _arr(v.colors, n * 3, np.float32) if v.colors else None,
            "uvs": _arr(v.uvs, n * 2, np.float32) if v.uvs else None,
            "indices": _arr(v.indices, int(v.index_count), np.uint32),
            "cellSize": float(v.cell_size),
        }
    finally:
        lib.vgd_proxy_free(handle)


def encode_vertex_buffer(raw: bytes, count: int, stride: int) -> bytes:
    """MeshoptEncoder.encodeVertexBufferLevel(raw, count, stride, 2, 1): a version-1 vertex-codec stream."""
    lib = _load()
    if len(raw) != count * stride:
        raise ValueError("raw size is not count x stride")
    cap = lib.vgd_vertex_encode_bound(count, stride)
    if cap == 0:
        raise ValueError(f"unsupported vertex size {stride}")
    out = ctypes.create_string_buffer(cap)
    src = ctypes.create_string_buffer(bytes(raw), max(len(raw), 1))
    n = lib.vgd_vertex_encode(out, cap, src, count, stride)
    if n == 0:
        raise ValueError("vertex encoding failed")
    return out.raw[:n]


def decode_vertex_buffer(blob: bytes, count: int, stride: int) -> bytes:
    """MeshoptDecoder.decodeVertexBuffer: `count` x `stride` bytes; raises ValueError with the codec's code."""
    lib = _load()
    out = ctypes.create_string_buffer(max(count * stride, 1))
    src = ctypes.create_string_buffer(bytes(blob), max(len(blob), 1))
    rc = lib.vgd_vertex_decode(out, count, stride, src, len(blob))
    if rc != 0:
        raise ValueError(f"vertex codec error {rc}")
    return out.raw[:count * stride]


def encode_index_buffer(indices) -> bytes:
    """MeshoptEncoder.encodeIndexBuffer(indices, count, 4): a version-1 index-codec stream."""
    lib = _load()
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1))
    cap = lib.vgd_index_encode_bound(len(idx), 0)
    out = ctypes.create_string_buffer(cap)
    n = lib.vgd_index_encode(out, cap, idx.ctypes.data, len(idx))
    if n == 0:
        raise ValueError("index encoding failed")
    return out.raw[:n]


def decode_index_buffer(blob: bytes, count: int) -> np.ndarray:
    """MeshoptDecoder.decodeIndexBuffer(target, count, 4, blob): uint32 indices (corners possibly rotated)."""
    lib = _load()
    out = np.zeros(max(count, 1), dtype=np.uint32)
    src = ctypes.create_string_buffer(bytes(blob), max(len(blob), 1))
    rc = lib.vgd_index_decode(out.ctypes.data, count, src, len(blob))
    if rc != 0:
        raise ValueError(f"index codec error {rc}")
    return out[:count]


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
