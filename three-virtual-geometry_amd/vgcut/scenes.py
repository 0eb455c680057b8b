"""Seeded synthetic scenes of the cut's benchmark configurations (inputs, not part of the cut).

icosphere_view  one icosphere of 1,003,520 triangles, one identity instance, one 1080p camera at 3R
ruins    procedural props (boxes, columns, arches, spheres) scattered uniformly over a square
world    props placed by Bridson Poisson-disk sampling over a noise heightfield, with per-mesh
         draw distances, and (terrain_grid > 0) the heightfield itself as one more mesh with one instance,
         so that drawing the previous frame's visible set gives ground occluders too
Every generator takes its sizes as arguments so tests can build small versions of the same scenes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import camera as cam
from . import scenegen


@dataclass
class SceneSpec:
    meshes: list  # VirtualMeshData
    matrices: List[np.ndarray]  # per mesh: (n, 16) float32 column-major
    max_draw_distance: List[float]
    min_pixel_radius: List[float] = field(default_factory=list)
    cast_shadow: List[bool] = field(default_factory=list)
    heightfield: Optional[dict] = None  # amplitude, frequency, seed of the ground (placement and occluder depth)
    props: Optional[int] = None  # meshes[:props] are the props (placed instances); later ones (the terrain) are not

    @property
    def instance_count(self) -> int:
        return int(sum(len(m) for m in self.matrices))

    @property
    def full_detail_triangles(self) -> int:
        return int(sum(d.leafTriangles * len(m) for d, m in zip(self.meshes, self.matrices)))


def yaw_scale_translate(yaw: np.ndarray, scale: np.ndarray, pos: np.ndarray) -> np.ndarray:
    """Column-major matrices of rotation about +y, uniform scale and translation."""
    n = len(yaw)
    m = np.zeros((n, 16), dtype=np.float64)
    c, s = np.cos(yaw) * scale, np.sin(yaw) * scale
    m[:, 0], m[:, 2] = c, -s
    m[:, 5] = scale
    m[:, 8], m[:, 10] = s, c
    m[:, 12:15] = pos
    m[:, 15] = 1.0
    return m.astype(np.float32)


def prop_params(rng: np.random.Generator, triangles: int, seed: int, kind: Optional[int] = None) -> dict:
    """Build parameters of a procedural prop of about `triangles` leaf triangles: box, column, arch or sphere."""
    if kind is None:
        kind = int(rng.integers(0, 4))
    if kind == 0:
        size = (rng.uniform(0.5, 2.5), rng.uniform(0.5, 2.5), rng.uniform(0.5, 2.5))
        return dict(kind=scenegen.BOX, seed=seed, size=size, noise_amplitude=0.05, noise_frequency=1.5, target_triangles=triangles)
    if kind == 1:
        size = (rng.uniform(0.3, 0.9), rng.uniform(3.0, 8.0), 0)
        return dict(kind=scenegen.COLUMN, seed=seed, size=size, noise_amplitude=0.05, noise_frequency=1.5, target_triangles=triangles)
    if kind == 2:
        size = (rng.uniform(3.0, 6.0), rng.uniform(2.0, 5.0), rng.uniform(0.5, 1.0))
        return dict(kind=scenegen.ARCH, seed=seed, size=size, noise_amplitude=0.05, noise_frequency=1.0, target_triangles=triangles)
    r = rng.uniform(0.8, 2.5)
    return dict(kind=scenegen.SPHERE, seed=seed, size=(r, 0, 0), noise_amplitude=0.05, noise_frequency=2.0 / r,
                target_triangles=triangles)


def prop_mesh(rng: np.random.Generator, triangles: int, seed: int, kind: Optional[int] = None):
    return scenegen.build_mesh(**prop_params(rng, triangles, seed, kind))


def build_props(rng: np.random.Generator, triangles, seed_base: int, threads: int = 16) -> list:
    """Builds one prop per entry of `triangles` (parameters drawn in order, DAGs built in parallel)."""
    from concurrent.futures import ThreadPoolExecutor

    params = [prop_params(rng, int(t), seed_base + i) for i, t in enumerate(triangles)]
    with ThreadPoolExecutor(max_workers=max(1, min(threads, len(params)))) as pool:
        return list(pool.map(lambda p: scenegen.build_mesh(**p), params))


def icosphere_view() -> Tuple[SceneSpec, cam.Camera, float]:
    """Plain single view: icosphere R = 1 of 1,003,520 triangles, camera at (0, 0, 3R), 1920 x 1080, 60 degrees."""
    data = scenegen.icosphere(triangles=1_003_520, radius=1.0, seed=1, amplitude=0.05)
    spec = SceneSpec([data], [np.eye(4, dtype=np.float32).reshape(1, 16)], [1e30], [0.0], [False])
    camera = cam.perspective_camera(60, 16 / 9, 0.01, 100).look_at((0, 0, 3), (0, 0, 0))
    return spec, camera, 1080.0


def log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n))


def ruins(mesh_count=48, tri_range=(200_000, 5_000_000), instances=113_106, half_extent=500.0, seed=2) -> SceneSpec:
    """Props with uniform positions over [-e, e]^2 on y = 0, yaw U[0, 2pi), scale U[0.5, 2]."""
    rng = np.random.default_rng(seed)
    tris = log_uniform(rng, tri_range[0], tri_range[1], mesh_count).astype(np.int64)
    meshes = build_props(rng, tris, seed * 1000)
    irng = np.random.default_rng(seed + 1)
    which = irng.integers(0, mesh_count, instances)
    pos = np.zeros((instances, 3))
    pos[:, 0] = irng.uniform(-half_extent, half_extent, instances)
    pos[:, 2] = irng.uniform(-half_extent, half_extent, instances)
    yaw = irng.uniform(0, 2 * math.pi, instances)
    scale = irng.uniform(0.5, 2.0, instances)
    all_m = yaw_scale_translate(yaw, scale, pos)
    mats = [all_m[which == k] for k in range(mesh_count)]
    return SceneSpec(meshes, mats, [1e30] * mesh_count, [0.0] * mesh_count, [False] * mesh_count,
                     heightfield=dict(amplitude=30.0, frequency=1 / 150.0, seed=seed + 3))


def ground_poses(spec: SceneSpec, count: int, half_extent: float, eye_height: float, seed: int):
    """Camera poses `eye_height` above the ground (y = 0 for the plane scenes), random yaw, slight pitch."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(count):
        x, z = rng.uniform(-0.8 * half_extent, 0.8 * half_extent, 2)
        ground = 0.0
        if spec.heightfield is not None and spec.heightfield.get("placement", False):
            ground = float(scenegen.heightfield(np.array([[x, z]], dtype=np.float32), spec.heightfield["amplitude"],
                                                spec.heightfield["frequency"], spec.heightfield["seed"])[0])
        yaw = rng.uniform(0, 2 * math.pi)
        pitch = rng.uniform(-0.15, 0.05)
        eye = np.array([x, ground + eye_height, z])
        d = np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
        poses.append((eye, eye + 100.0 * d))
    return poses


def terrain_mesh(size: float, amplitude: float, frequency: float, seed: int, grid: int):
    """The noise heightfield over [-size/2, size/2]^2 as a grid of grid x grid quads, DAG-built like the props
    (two triangles per quad, facing +y)."""
    from . import dag

    xs = np.linspace(-size / 2, size / 2, grid + 1, dtype=np.float64)
    gx, gz = np.meshgrid(xs, xs, indexing="ij")  # vertex (i, j) at x = xs[i], z = xs[j]
    xz = np.stack([gx.ravel(), gz.ravel()], 1).astype(np.float32)
    y = scenegen.heightfield(xz, amplitude, frequency, seed)
    pos = np.stack([xz[:, 0], y, xz[:, 1]], 1).astype(np.float32)
    i, j = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    v00 = (i * (grid + 1) + j).ravel()
    v01, v10, v11 = v00 + 1, v00 + grid + 1, v00 + grid + 2
    idx = np.stack([v00, v01, v11, v00, v11, v10], 1).astype(np.uint32).ravel()
    return dag.build_virtual_mesh(pos, idx)


def world(mesh_count=64, tri_range=(100_000, 2_000_000), instances=3_000_000, size=6144.0, amplitude=200.0,
          draw_range=(200.0, 6000.0), seed=6, meshes=None, terrain_grid=0) -> SceneSpec:
    """Props by Bridson Poisson-disk sampling over a size x size noise heightfield (y = terrain height).
    `meshes`: a prop set to reuse (built from `seed` otherwise). `terrain_grid` > 0 adds the heightfield
    as one more mesh (terrain_grid^2 quads, one identity instance, no draw distance)."""
    rng = np.random.default_rng(seed)
    tris = log_uniform(rng, tri_range[0], tri_range[1], mesh_count).astype(np.int64)
    if meshes is None:
        meshes = build_props(rng, tris, seed * 1000)
    mesh_count = len(meshes)
    # Disk radius for about 1.1x the wanted count (Bridson with 12 candidates packs about 0.595 points per
    # r^2), then an even random subsample to exactly `instances`.
    radius = math.sqrt(0.595 * size * size / (instances * 1.1))
    xz = scenegen.poisson_disk(-size / 2, -size / 2, size, size, radius, seed + 1, candidates=12,
                               max_points=int(instances * 1.6) + 16)
    prng = np.random.default_rng(seed + 2)
    if len(xz) > instances:
        xz = xz[np.sort(prng.choice(len(xz), instances, replace=False))]
    n = len(xz)
    hf = dict(amplitude=amplitude, frequency=1.0 / 900.0, seed=seed + 3, placement=True)
    y = scenegen.heightfield(xz, hf["amplitude"], hf["frequency"], hf["seed"])
    pos = np.stack([xz[:, 0].astype(np.float64), y.astype(np.float64), xz[:, 1].astype(np.float64)], axis=1)
    which = prng.integers(0, mesh_count, n)
    yaw = prng.uniform(0, 2 * math.pi, n)
    scale = prng.uniform(0.5, 2.0, n)
    all_m = yaw_scale_translate(yaw, scale, pos)
    mats = [all_m[which == k] for k in range(mesh_count)]
    drng = np.random.default_rng(seed + 2000)
    draw = list(drng.uniform(draw_range[0], draw_range[1], mesh_count))
    spec = SceneSpec(list(meshes), mats, draw, [0.0] * mesh_count, [False] * mesh_count, heightfield=hf, props=mesh_count)
    if terrain_grid > 0:
        spec.meshes.append(terrain_mesh(size, amplitude, hf["frequency"], hf["seed"], terrain_grid))
        spec.matrices.append(np.eye(4, dtype=np.float32).reshape(1, 16))
        spec.max_draw_distance.append(1e30)
        spec.min_pixel_radius.append(0.0)
        spec.cast_shadow.append(False)
    return spec


def flythrough(spec: SceneSpec, frames: int, size: float, height: float, seed: int):
    """A seeded Catmull-Rom path `height` above the ground, looking along the path."""
    rng = np.random.default_rng(seed)
    ctrl = rng.uniform(-0.35 * size, 0.35 * size, (6, 2))
    pts = []
    for f in range(frames + 1):
        u = f / frames * (len(ctrl) - 3)
        i = min(int(u), len(ctrl) - 4)
        t = u - i
        p0, p1, p2, p3 = ctrl[i:i + 4]
        q = 0.5 * ((2 * p1) + (-p0 + p2) * t + (2 * p0 - 5 * p1 + 4 * p2 - p3) * t * t + (-p0 + 3 * p1 - 3 * p2 + p3) * t ** 3)
        pts.append(q)
    pts = np.array(pts)
    hf = spec.heightfield
    g = scenegen.heightfield(pts.astype(np.float32), hf["amplitude"], hf["frequency"], hf["seed"]) if hf else np.zeros(len(pts))
    poses = []
    for f in range(frames):
        eye = np.array([pts[f, 0], g[f] + height, pts[f, 1]])
        ahead = np.array([pts[f + 1, 0], g[f] + height - 0.1 * height, pts[f + 1, 1]])
        poses.append((eye, ahead))
    return poses


def load_into(scene, spec: SceneSpec, occlusion: bool = False):
    """Adds a spec's meshes, instances and flags to a vgcut native.Scene; returns the mesh slots."""
    slots = []
    for k, (data, mats) in enumerate(zip(spec.meshes, spec.matrices)):
        if len(mats) == 0:
            continue
        slot = scene.add_mesh(data)
        scene.set_instances(slot, mats)
        scene.set_flags(slot, spec.max_draw_distance[k], spec.min_pixel_radius[k] if spec.min_pixel_radius else 0.0, True,
                        spec.cast_shadow[k] if spec.cast_shadow else False, occlusion)
        slots.append(slot)
    return slots
