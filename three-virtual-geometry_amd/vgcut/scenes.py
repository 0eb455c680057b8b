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


def mixed_views(spec: SceneSpec, count: int = 64, size: float = 6144.0):
    """The batched workload's mixed views of a world scene, as (vg_view, occluder depth or None), slots 0..count-1:
    4 main cameras on the fly-through path (2 at 1920 x 1080, 2 at 3840 x 2160, device pixels, occlusion
    against the terrain's depth at the reference's HZB sizes), 4 orthographic shadow cascades per main
    camera at 2048^2 (half extents 20-540 m, toward a fixed sun), 6 cube-map faces of a probe near each
    main camera at 512^2 (90 degrees, 60 m far) and the rest 720p viewers at seeded poses (occlusion
    against the terrain at 640 x 360). Orthographic views keep the reference's perspective LOD formula
    and every instance however small on screen (minPixelRadius 0)."""
    from .virtual_geometry import make_view

    hf = spec.heightfield
    path = flythrough(spec, 16, size, 30.0, seed=9)
    views = []
    rng = np.random.default_rng(10)
    sun = np.array([-0.4, -1.0, -0.3])
    sun /= np.linalg.norm(sun)

    def terrain_depth(camera, w, h):
        vp = cam.mat_mul(camera.projectionMatrix, camera.matrixWorldInverse)
        return scenegen.heightfield_depth(vp, w, h, (0.0, 0.0), size / 2, 256, hf["amplitude"], hf["frequency"], hf["seed"])

    mains = []
    for i in range(4):
        eye, target = path[i * 4]
        width, height = (1920, 1080) if i < 2 else (3840, 2160)
        camera = cam.perspective_camera(60, width / height, 0.1, 20000).look_at(eye, target)
        mains.append((camera, eye, target))
        hz = (960, 540) if i < 2 else (1024, 576)
        views.append((camera, height, True, terrain_depth(camera, *hz)))
    for camera, eye, target in mains:  # 4 cascades per main camera, centred ahead of it
        fwd = np.asarray(target, dtype=np.float64) - np.asarray(eye, dtype=np.float64)
        fwd /= np.linalg.norm(fwd)
        for half in (20.0, 60.0, 180.0, 540.0):
            center = np.asarray(eye) + fwd * half
            ortho = cam.orthographic_camera(-half, half, half, -half, 0.1, 2000).look_at(center - sun * (half + 300), center)
            views.append((ortho, 2048, False, None))
    for camera, eye, target in mains:  # a reflection probe near each main camera: 6 faces
        probe = np.asarray(eye) + np.array([5.0, -10.0, 5.0])
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            up = (0, 0, 1) if abs(d[1]) == 1 else (0, 1, 0)
            face = cam.perspective_camera(90, 1.0, 0.1, 60.0).look_at(probe, probe + np.array(d, dtype=np.float64), up)
            views.append((face, 512, False, None))
    xz = rng.uniform(-0.45 * size, 0.45 * size, (count, 2)).astype(np.float32)
    g = scenegen.heightfield(xz, hf["amplitude"], hf["frequency"], hf["seed"])
    k = 0
    while len(views) < count:
        yaw = rng.uniform(0, 2 * math.pi)
        eye = np.array([xz[k, 0], g[k] + rng.uniform(2, 60), xz[k, 1]])
        target = eye + np.array([math.sin(yaw), -0.1, math.cos(yaw)]) * 100
        camera = cam.perspective_camera(60, 16 / 9, 0.1, 20000).look_at(eye, target)
        views.append((camera, 720, True, terrain_depth(camera, 640, 360)))
        k += 1
    out = []
    for slot, (camera, height, occl, depth) in enumerate(views[:count]):
        min_px = 0.7 if camera.isPerspectiveCamera else 0.0
        out.append((make_view(camera, height, slot=slot, occlusion=occl, min_pixel_radius=min_px), depth))
    return out


# Instance counts of the served resident scenes.
SERVE_SIZES = (12_000, 50_000, 113_000, 152_000, 500_000, 1_000_000, 2_000_000, 3_000_000)


def serving_specs(props, sizes=SERVE_SIZES):
    """The served resident scenes: world scenes of `sizes` instances over the prop set `props`, each on a
    square sized for the world's instance density, seeded 100, 101, ...; returns (spec, side length)."""
    out = []
    for k, n in enumerate(sizes):
        side = 6144.0 * math.sqrt(n / 3_000_000)
        out.append((world(instances=n, size=side, seed=100 + k, meshes=props), side))
    return out


def serve_requests(resident, count: int, seed: int = 11):
    """A seeded pool of cut requests over `resident` [(spec, side)]: (scene index, ctypes vg_view array, view
    count). A request is a scene, 1-16 ground views at 720p or 1080p, an error threshold in [0.5, 2] px and,
    for 30% of them, a triangle budget in [2M, 10M] (threshold range 0.5-3 px)."""
    from . import native
    from .virtual_geometry import make_view

    rng = np.random.default_rng(seed)
    pool = []
    for _ in range(count):
        k = int(rng.integers(0, len(resident)))
        spec, side = resident[k]
        hf = spec.heightfield
        n = int(rng.integers(1, 17))
        t = float(rng.uniform(0.5, 2.0))
        budget = int(rng.integers(2_000_000, 10_000_001)) if rng.random() < 0.3 else 0
        xz = rng.uniform(-0.45 * side, 0.45 * side, (n, 2)).astype(np.float32)
        g = scenegen.heightfield(xz, hf["amplitude"], hf["frequency"], hf["seed"])
        arr = (native.View * n)()
        for i in range(n):
            yaw = rng.uniform(0, 2 * math.pi)
            eye = np.array([xz[i, 0], g[i] + rng.uniform(2, 60), xz[i, 1]])
            target = eye + np.array([math.sin(yaw), -0.1, math.cos(yaw)]) * 100
            height = int(rng.choice([720, 1080]))
            camera = cam.perspective_camera(60, 16 / 9, 0.1, 20000).look_at(eye, target)
            arr[i] = make_view(camera, height, slot=i, error_threshold=t, triangle_budget=budget, threshold_range=(0.5, 3.0))
        pool.append((k, arr, n))
    return pool
