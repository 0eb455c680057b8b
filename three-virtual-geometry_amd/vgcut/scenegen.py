"""Python access to the synthetic scene producer (csrc/vgscene.cpp).

Builds meshlet LOD DAGs in the reference's ``VirtualMeshData`` layout
(src/core/preprocess/buildVirtualMesh.ts:75-138) from seeded procedural meshes, places instances
and renders synthetic occluder depth. These are inputs of the cut, not part of it.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Optional

import numpy as np

from . import _native

SPHERE, BOX, COLUMN, ARCH, TERRAIN = 0, 1, 2, 3, 4


class _Params(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("seed", ctypes.c_uint32),
        ("size", ctypes.c_double * 3),
        ("noise_amplitude", ctypes.c_double),
        ("noise_frequency", ctypes.c_double),
        ("resolution", ctypes.c_uint32),
        ("target_triangles", ctypes.c_uint32),
    ]


class _Triangles(ctypes.Structure):
    _fields_ = [
        ("vertex_count", ctypes.c_uint32),
        ("positions", ctypes.POINTER(ctypes.c_float)),
        ("index_count", ctypes.c_uint64),
        ("indices", ctypes.POINTER(ctypes.c_uint32)),
    ]


class _View(ctypes.Structure):
    _fields_ = [
        ("meshlet_count", ctypes.c_uint32),
        ("meshlet_bounds", ctypes.POINTER(ctypes.c_float)),
        ("meshlet_info", ctypes.POINTER(ctypes.c_uint32)),
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
        ("meshlet_triangles", ctypes.POINTER(ctypes.c_uint32)),
        ("triangle_count", ctypes.c_uint64),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(_native.library_path("libvgscene.so"))
        lib.vgs_mesh_build.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(ctypes.c_void_p)]
        lib.vgs_mesh_build.restype = ctypes.c_int
        lib.vgs_mesh_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(_View)]
        lib.vgs_mesh_get.restype = None
        lib.vgs_mesh_free.argtypes = [ctypes.c_void_p]
        lib.vgs_mesh_free.restype = None
        lib.vgs_mesh_triangles.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_Triangles)]
        lib.vgs_mesh_triangles.restype = ctypes.c_int
        lib.vgs_triangles_free.argtypes = [ctypes.POINTER(_Triangles)]
        lib.vgs_triangles_free.restype = None
        lib.vgs_poisson_disk.argtypes = [ctypes.c_double] * 5 + [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        lib.vgs_poisson_disk.restype = ctypes.c_uint32
        lib.vgs_heightfield.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p]
        lib.vgs_heightfield.restype = None
        lib.vgs_heightfield_depth.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32] + [ctypes.c_double] * 3 + [
            ctypes.c_uint32, ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p]
        lib.vgs_heightfield_depth.restype = None
        _lib = lib
    return _lib


@dataclasses.dataclass
class VirtualMeshData:
    """The fields of the reference's ``VirtualMeshData`` the cut reads, plus the DAG links.

    Mirrors src/core/preprocess/buildVirtualMesh.ts:75-138. ``meshletTriangles`` (the local vertex ids
    the expand pass writes as indices) is always produced; positions, the flat index list and meshlet
    vertex lists come from the DAG builder (``vgcut.dag``) only, and with them the normals, colors and uvs
    the builder was given (extended over the vertices its levels add).
    """

    meshletCount: int
    meshletBounds: np.ndarray  # float32 [n * 16]
    meshletInfo: np.ndarray  # uint32 [n * 4]
    levelRanges: np.ndarray  # uint32 [levels * 2]
    levelErrors: np.ndarray  # float32 [levels * 2]
    levelGuardErrors: np.ndarray  # float32 [levels * 2]
    replacementGroup: np.ndarray  # int32 [n]
    replacementStart: np.ndarray  # uint32 [groups + 1]
    replacementIndices: np.ndarray  # uint32
    boundingSphere: np.ndarray  # float32 [4]
    lodBoundsSphere: np.ndarray  # float32 [4]
    cullBoundsSphere: np.ndarray  # float32 [4]
    leafTriangles: int
    leafMeshlets: int
    rootTriangles: int
    rootMeshlets: int
    meshletTriangles: Optional[np.ndarray] = None  # uint32 per triangle: local ids a | b << 8 | c << 16
    # Produced by the DAG builder (vgcut.dag) only: the vertex buffer, the flat meshlet index lists and the
    # per-meshlet vertex lists that meshletTriangles' local ids index (meshletInfo[3] is the offset).
    positions: Optional[np.ndarray] = None  # float32 [vertices * 3]
    indices: Optional[np.ndarray] = None  # uint32, meshletInfo[1] is each meshlet's first
    meshletVertices: Optional[np.ndarray] = None  # uint32
    lodLevels: int = 0
    normals: Optional[np.ndarray] = None  # float32 [vertices * 3]
    colors: Optional[np.ndarray] = None  # float32 [vertices * 3]
    uvs: Optional[np.ndarray] = None  # float32 [vertices * 2]
    voxelLevels: int = 0  # stats.voxelLevels
    aggregateGroups: int = 0  # stats.aggregateGroups

    @property
    def levelCount(self) -> int:
        return len(self.levelRanges) // 2


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _params(kind, seed, size, noise_amplitude, noise_frequency, resolution, target_triangles) -> _Params:
    p = _Params()
    p.kind = kind
    p.seed = seed
    for i in range(3):
        p.size[i] = float(size[i])
    p.noise_amplitude = noise_amplitude
    p.noise_frequency = noise_frequency
    p.resolution = resolution
    p.target_triangles = target_triangles
    return p


def triangle_mesh(kind: int, seed: int = 1, size=(1.0, 1.0, 1.0), noise_amplitude: float = 0.05, noise_frequency: float = 4.0,
                  resolution: int = 16, target_triangles: int = 0):
    """The procedural mesh as one watertight indexed triangle mesh: (float32 [v, 3] positions, uint32 indices)."""
    lib = _load()
    p = _params(kind, seed, size, noise_amplitude, noise_frequency, resolution, target_triangles)
    t = _Triangles()
    rc = lib.vgs_mesh_triangles(ctypes.byref(p), ctypes.byref(t))
    if rc != 0:
        raise ValueError(f"vgs_mesh_triangles failed ({rc}) for kind {kind}")
    try:
        pos = _arr(t.positions, t.vertex_count * 3, np.float32).reshape(-1, 3)
        idx = _arr(t.indices, int(t.index_count), np.uint32)
    finally:
        lib.vgs_triangles_free(ctypes.byref(t))
    return pos, idx


def build_mesh(kind: int, seed: int = 1, size=(1.0, 1.0, 1.0), noise_amplitude: float = 0.05,
               noise_frequency: float = 4.0, resolution: int = 16, target_triangles: int = 0, builder: str = "dag",
               threads: int = 0) -> VirtualMeshData:
    """Builds one procedural mesh and its DAG.

    ``builder="dag"`` tessellates the mesh (``triangle_mesh``) and runs the DAG builder (vgcut.dag, the
    reference's buildVirtualMesh at its defaults) with ``threads`` workers; ``builder="patch"`` uses the
    grid-patch construction of csrc/vgscene.cpp, which builds in a fraction of the time.
    """
    if builder == "dag":
        from . import dag

        pos, idx = triangle_mesh(kind, seed, size, noise_amplitude, noise_frequency, resolution, target_triangles)
        return dag.build_virtual_mesh(pos, idx, threads=threads)
    if builder != "patch":
        raise ValueError(f"unknown builder {builder}")
    lib = _load()
    p = _params(kind, seed, size, noise_amplitude, noise_frequency, resolution, target_triangles)
    handle = ctypes.c_void_p()
    rc = lib.vgs_mesh_build(ctypes.byref(p), ctypes.byref(handle))
    if rc != 0:
        raise ValueError(f"vgs_mesh_build failed ({rc}) for kind {kind}")
    try:
        v = _View()
        lib.vgs_mesh_get(handle, ctypes.byref(v))
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
            meshletTriangles=_arr(v.meshlet_triangles, int(v.triangle_count), np.uint32),
        )
    finally:
        lib.vgs_mesh_free(handle)


def icosphere(triangles: int = 1_003_520, radius: float = 1.0, seed: int = 1, amplitude: float = 0.05) -> VirtualMeshData:
    """Seeded icosphere of about ``triangles`` leaf triangles, radially displaced by 3-octave value noise."""
    return build_mesh(SPHERE, seed=seed, size=(radius, 0, 0), noise_amplitude=amplitude, noise_frequency=4.0 / radius,
                      target_triangles=triangles)


def reference_terrain() -> VirtualMeshData:
    """The reference tests' sine terrain (tests/cut.test.ts:14-19): 400 x 400, 96 x 96 quads."""
    return build_mesh(TERRAIN, resolution=96)


def poisson_disk(x0: float, z0: float, w: float, h: float, radius: float, seed: int, candidates: int = 12,
                 max_points: Optional[int] = None) -> np.ndarray:
    lib = _load()
    if max_points is None:
        max_points = int(w * h / (radius * radius) * 1.2) + 16
    out = np.empty((max_points, 2), dtype=np.float32)
    n = lib.vgs_poisson_disk(x0, z0, w, h, radius, seed, candidates, max_points, out.ctypes.data)
    return out[:n].copy()


def heightfield(xz: np.ndarray, amplitude: float, frequency: float, seed: int) -> np.ndarray:
    lib = _load()
    xz = np.ascontiguousarray(xz, dtype=np.float32)
    out = np.empty(len(xz), dtype=np.float32)
    lib.vgs_heightfield(xz.ctypes.data, len(xz), amplitude, frequency, seed, out.ctypes.data)
    return out


def heightfield_depth(view_proj: np.ndarray, width: int, height: int, center_xz, extent: float, grid: int,
                      amplitude: float, frequency: float, seed: int) -> np.ndarray:
    """WebGPU depth ([0, 1], far = 1) of a rasterized noise heightfield, row 0 at the top."""
    lib = _load()
    vp = np.ascontiguousarray(np.asarray(view_proj, dtype=np.float32).reshape(16))
    out = np.empty(width * height, dtype=np.float32)
    lib.vgs_heightfield_depth(vp.ctypes.data, width, height, float(center_xz[0]), float(center_xz[1]), extent, grid,
                              amplitude, frequency, seed, out.ctypes.data)
    return out.reshape(height, width)


def library_exists() -> bool:
    return os.path.exists(_native.library_path("libvgscene.so"))
