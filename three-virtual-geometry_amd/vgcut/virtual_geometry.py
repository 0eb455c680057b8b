"""Host side of the cut, mirroring the reference's ``VirtualGeometry`` / ``VirtualMesh`` surface
(src/core/runtime/VirtualGeometry.ts:62-633, VirtualMesh.ts:39-331) over the HIP library.

Kept: the levers (errorThreshold, lodBlend, maxErrorThreshold, triangleBudget, thresholdRange,
minPixelRadius, frustumCulling, shadowErrorScale, shadowSweep), ``createMesh`` with per-mesh
``maxDrawDistance`` / ``minPixelRadius``, ``update(camera)``, ``readStats()`` and the CPU budget glide
controller (adaptThreshold / glideTo / stepThreshold). Raster, materials, the scene-graph hooks and
the occluder depth pass are not part of this path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import camera as cam
from . import native

THRESHOLD_RATE_UP = 1.03  # VirtualGeometry.ts:21-22
THRESHOLD_RATE_DOWN = 1.01
NO_DRAW_DISTANCE = 1e30  # GeometryPool.ts:111
STATE_FIXED = 10000  # GeometryPool.ts:131
S_SCALE, S_FILL = 0, 2
FILL_HIGH, FILL_LOW = 0.9, 0.75  # GeometryPool.ts:136-137


def refine_threshold(threshold: float, use: float) -> float:
    """VirtualGeometry.ts:60."""
    return threshold * math.pow(use / 0.82, 0.25)


def make_view(camera: cam.Camera, viewport_height: float, *, slot: int = 0, error_threshold: float = 1.0,
              shadow_error_scale: float = 3.0, lod_blend: float = 2.0, max_error_threshold: float = 64.0,
              min_pixel_radius: float = 0.7, frustum_culling: bool = True, occlusion: bool = False,
              shadow_sweep=(0.0, 0.0, 0.0), triangle_budget: int = 0, threshold_range=(0.5, 3.0)) -> native.View:
    """The per-view uniforms VirtualGeometry.update() sets (VirtualGeometry.ts:400-437), as a vg_view."""
    v = native.View()
    view = camera.matrixWorldInverse
    proj = camera.projectionMatrix
    planes = cam.frustum_planes(cam.mat_mul(proj, view))
    for i in range(16):
        v.view_matrix[i] = float(view[i])
        v.projection_matrix[i] = float(proj[i])
    for i in range(24):
        v.frustum_planes[i] = float(planes.reshape(24)[i])
    for i in range(3):
        v.shadow_sweep[i] = float(shadow_sweep[i])
    v.proj_scale = cam.proj_scale(camera, viewport_height)
    v.near_plane = camera.near
    v.error_threshold = error_threshold
    v.shadow_error_scale = shadow_error_scale
    v.lod_blend_max = max(1.0, lod_blend)
    v.lod_scale_limit = max(1.0, max_error_threshold / error_threshold)
    v.min_pixel_radius = min_pixel_radius
    v.frustum_culling = 1 if frustum_culling else 0
    # The occlusion test assumes a perspective camera (OcclusionCulling.ts:232-237).
    # occlusion=2: two-phase (the cut records its occlusion rejects for vg_cut_phase2).
    v.occlusion = (2 if occlusion == 2 else 1) if (occlusion and camera.isPerspectiveCamera) else 0
    v.slot = slot
    v.triangle_budget = int(triangle_budget)
    v.threshold_range[0] = threshold_range[0]
    v.threshold_range[1] = threshold_range[1]
    v.max_error_threshold = max_error_threshold
    return v


@dataclass
class VirtualGeometryStats:
    """VirtualGeometry.ts:24-47."""

    drawnMeshlets: int = 0
    drawnTriangles: int = 0
    fullDetailTriangles: int = 0
    instances: int = 0
    visibleInstances: int = 0
    requestedTriangles: int = 0
    overflow: bool = False
    capacityUse: float = 0.0
    occludedInstances: int = 0
    shadowTriangles: int = 0
    thresholdScale: float = 1.0


class VirtualMesh:
    """Instances of one DAG in the scene (VirtualMesh.ts:91-331, cut-side levers only)."""

    def __init__(self, context: "VirtualGeometry", data, matrices: np.ndarray, colors: Optional[np.ndarray] = None,
                 maxDrawDistance: Optional[float] = None, minPixelRadius: Optional[float] = None):
        self.context = context
        self.data = data
        m = np.ascontiguousarray(matrices, dtype=np.float32).reshape(-1, 16)
        self.instanceCount = len(m)
        self.fullDetailTriangles = data.leafTriangles * self.instanceCount
        self.slot = context.scene.add_mesh(data)
        self.gpuToInput = context.scene.set_instances(self.slot, m, colors)
        self.gpuIndexOf = np.empty_like(self.gpuToInput)
        self.gpuIndexOf[self.gpuToInput] = np.arange(self.instanceCount, dtype=np.uint32)
        self.instanceBase, _ = context.scene.instance_base(self.slot)
        self._max_draw = NO_DRAW_DISTANCE
        self._min_px = 0.0
        self.visible = True
        self.castShadow = False
        self.occlusionCulling = True
        if maxDrawDistance is not None:
            self.maxDrawDistance = maxDrawDistance
        if minPixelRadius is not None:
            self.minPixelRadius = minPixelRadius
        self.sync_flags()

    @property
    def maxDrawDistance(self) -> float:
        return math.inf if self._max_draw >= NO_DRAW_DISTANCE else self._max_draw

    @maxDrawDistance.setter
    def maxDrawDistance(self, d: float):
        self._max_draw = max(0.0, d) if math.isfinite(d) else NO_DRAW_DISTANCE  # VirtualMesh.ts:142-144
        self.sync_flags()

    @property
    def minPixelRadius(self) -> float:
        return self._min_px

    @minPixelRadius.setter
    def minPixelRadius(self, px: float):
        self._min_px = max(0.0, px)  # VirtualMesh.ts:150-152
        self.sync_flags()

    def sync_flags(self):
        self.context.scene.set_flags(self.slot, self._max_draw, self._min_px, self.visible, self.castShadow,
                                     self.occlusionCulling and self.context.occlusion)


class VirtualGeometry:
    """Context: capacities, levers, per-frame update, stats and the budget controller."""

    def __init__(self, maxTriangles: int = 10_000_000, maxShadowTriangles: int = 4_000_000, maxMeshlets: int = 1 << 20,
                 maxViews: int = 1, meshSlots: int = 64, hzbMaxSize: int = 1024, device: int = 0):
        self.ctx = native.Context(maxTriangles, maxShadowTriangles, min(maxMeshlets, 1 << 24), maxViews, meshSlots,
                                  hzbMaxSize, 256, device)
        self.scene = native.Scene(self.ctx)
        self.capacity = dict(triangles=maxTriangles, shadowTriangles=maxShadowTriangles, meshlets=min(maxMeshlets, 1 << 24))
        self.meshes: List[VirtualMesh] = []
        self.errorThreshold = 1.0
        self.lodBlend = 2.0
        self.maxErrorThreshold = 64.0
        self.triangleBudget = 0
        self.thresholdRange = [0.5, 3.0]
        self.minPixelRadius = 0.7
        self.frustumCulling = True
        self.shadowErrorScale = 3.0
        self.shadowSweep = (0.0, 0.0, 0.0)
        self.occlusion = False
        self.lastStats: Optional[VirtualGeometryStats] = None
        self.frameCount = 0
        self._out_of_band = 0
        self._threshold_target: Optional[float] = None
        self._threshold_written = math.nan
        self._stats_threshold = 1.0

    # ------------------------------------------------------------------ scene
    def createMesh(self, data, instances: dict, options: Optional[dict] = None) -> VirtualMesh:
        """VirtualGeometry.createMesh (VirtualGeometry.ts:231-233): instances = {matrices, colors?}."""
        options = options or {}
        mesh = VirtualMesh(self, data, instances["matrices"], instances.get("colors"), options.get("maxDrawDistance"),
                           options.get("minPixelRadius"))
        self.meshes.append(mesh)
        return mesh

    def set_hzb(self, depth: np.ndarray, view_slot: int = 0):
        """Depth of last frame's occluders for this view (OcclusionCulling.beginFrame + HZB build)."""
        self.scene.set_hzb(view_slot, depth)

    # ------------------------------------------------------------------ per frame
    def view_for(self, camera: cam.Camera, viewport_height: float, slot: int = 0) -> native.View:
        return make_view(camera, viewport_height, slot=slot, error_threshold=self.errorThreshold,
                         shadow_error_scale=self.shadowErrorScale, lod_blend=self.lodBlend,
                         max_error_threshold=self.maxErrorThreshold, min_pixel_radius=self.minPixelRadius,
                         frustum_culling=self.frustumCulling, occlusion=self.occlusion, shadow_sweep=self.shadowSweep)

    def update(self, camera: cam.Camera, viewport_height: float, slot: int = 0, wait: bool = True):
        """VirtualGeometry.update (VirtualGeometry.ts:392-507) for one camera: glide, uniforms, cut."""
        self.step_threshold()
        for mesh in self.meshes:
            mesh.sync_flags()
        self.scene.cut([self.view_for(camera, viewport_height, slot)])
        self.frameCount += 1
        if wait:
            self.scene.sync()
        if self.triangleBudget > 0:
            self._stats_threshold = self.errorThreshold
            stats = self.readStats(slot)
            self.adapt_threshold(stats)

    def update_views(self, views: Sequence[native.View], graph: bool = False):
        """Batched form: several views in one launch sequence (one slot each)."""
        self.scene.cut(list(views), graph=graph)
        self.scene.sync()

    # ------------------------------------------------------------------ stats and the budget glide
    def readStats(self, slot: int = 0) -> VirtualGeometryStats:
        """VirtualGeometry.readStats over one pool (VirtualGeometry.ts:591-632, GeometryPool.ts:1135-1151)."""
        counters, _, _, _ = self.scene.read(slot, max_entries=1)
        words = counters
        s = VirtualGeometryStats()
        for mesh in self.meshes:
            if mesh.visible:
                s.fullDetailTriangles += mesh.fullDetailTriangles
                s.instances += mesh.instanceCount
        cap = self.capacity
        base = len(words) - 8
        scale = [words[base + S_SCALE] / STATE_FIXED, words[base + S_SCALE + 1] / STATE_FIXED]

        def cut(c):
            return dict(selected=int(words[2 + c * 5]), drawn=int(words[2 + c * 5 + 1]), survived=int(words[2 + c * 5 + 2]),
                        requestedTriangles=words[2 + c * 5 + 4] / 3)

        cam0, sh = cut(0), cut(1)
        s.drawnMeshlets = cam0["drawn"]
        s.drawnTriangles = min(cam0["requestedTriangles"], cap["triangles"])
        s.requestedTriangles = cam0["requestedTriangles"]
        s.visibleInstances = cam0["survived"]
        s.occludedInstances = int(words[1])
        s.shadowTriangles = min(sh["requestedTriangles"], cap["shadowTriangles"])
        s.thresholdScale = max(1.0, scale[0])
        s.overflow = (cam0["selected"] > cap["meshlets"] or cam0["requestedTriangles"] > cap["triangles"]
                      or sh["selected"] > cap["meshlets"] or sh["requestedTriangles"] > cap["shadowTriangles"])
        s.capacityUse = max(cam0["requestedTriangles"] / cap["triangles"], cam0["selected"] / cap["meshlets"],
                            sh["requestedTriangles"] / cap["shadowTriangles"], sh["selected"] / cap["meshlets"])
        self.lastStats = s
        return s

    def adapt_threshold(self, stats: VirtualGeometryStats):
        """adaptThreshold (VirtualGeometry.ts:549-568)."""
        t = self._stats_threshold
        now = self.errorThreshold
        use = stats.capacityUse
        if self._threshold_target is not None and self._threshold_target < now and use >= 0.75:
            self._threshold_target = None
        if stats.drawnTriangles <= 0:
            return
        ratio = stats.drawnTriangles / self.triangleBudget
        if abs(ratio - 1) < 0.15:
            self._out_of_band = 0
            return
        self._out_of_band += 1
        if self._out_of_band < 2:
            return
        lo, hi = self.thresholdRange
        nxt = min(hi, max(lo, t * math.sqrt(ratio)))
        if nxt < t:
            nxt = max(nxt, refine_threshold(t, use))
        self.glide_to(nxt)

    def glide_to(self, target: float):
        """glideTo (VirtualGeometry.ts:570-572)."""
        self._threshold_target = None if target == self.errorThreshold else target

    def step_threshold(self):
        """stepThreshold (VirtualGeometry.ts:575-588)."""
        t = self.errorThreshold
        if t != self._threshold_written and not math.isnan(self._threshold_written):
            self._threshold_target = None
        target = self._threshold_target
        if target is not None:
            nxt = min(target, t * THRESHOLD_RATE_UP) if target > t else max(target, t / THRESHOLD_RATE_DOWN)
            self.errorThreshold = nxt
            if nxt == target:
                self._threshold_target = None
        self._threshold_written = self.errorThreshold

    def close(self):
        self.scene.close()
        self.ctx.close()
