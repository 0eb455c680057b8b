"""Camera math in three.js conventions (fp64, column-major 16-element matrices).

VirtualGeometry.update() (src/core/runtime/VirtualGeometry.ts:414-437) reads the camera's
matrixWorldInverse and projection matrix, derives projScale = P[5] * height / 2 and extracts the
frustum planes with Frustum.setFromProjectionMatrix(P * V, WebGPUCoordinateSystem). These are
restated here from three.js's published formulas (Matrix4.makePerspective / makeOrthographic with
the WebGPU depth range [0, 1], Matrix4.lookAt, Frustum.setFromProjectionMatrix). three.js itself is
not available, so bit-equality with its fp64 results is not pinned.
"""
from __future__ import annotations

import math

import numpy as np


def _normalize(v):
    n = math.sqrt(float(np.dot(v, v)))
    return v / n if n > 0 else v


def look_at_world(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """World matrix of a camera at `eye` looking at `target` (its -z axis points at the target)."""
    eye = np.asarray(eye, dtype=np.float64)
    z = _normalize(eye - np.asarray(target, dtype=np.float64))
    if float(np.dot(z, z)) == 0:
        z = np.array([0.0, 0.0, 1.0])
    upv = np.asarray(up, dtype=np.float64)
    x = np.cross(upv, z)
    if float(np.dot(x, x)) == 0:
        # up parallel to the view direction: nudge as three.js does
        z2 = z.copy()
        if abs(upv[2]) == 1:
            z2[0] += 0.0001
        else:
            z2[2] += 0.0001
        z = _normalize(z2)
        x = np.cross(upv, z)
    x = _normalize(x)
    y = np.cross(z, x)
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = x, y, z, eye
    return m.T.reshape(16).copy()  # column-major flat


def invert_rigid(world: np.ndarray) -> np.ndarray:
    m = np.asarray(world, dtype=np.float64).reshape(4, 4).T  # row-major
    r = m[:3, :3]
    t = m[:3, 3]
    inv = np.eye(4)
    inv[:3, :3] = r.T
    inv[:3, 3] = -r.T @ t
    return inv.T.reshape(16).copy()


def perspective(fov_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """PerspectiveCamera.updateProjectionMatrix + makePerspective, WebGPU depth range."""
    top = near * math.tan(math.radians(0.5 * fov_deg))
    height = 2 * top
    width = aspect * height
    left = -0.5 * width
    right, bottom = left + width, top - height
    x = 2 * near / (right - left)
    y = 2 * near / (top - bottom)
    a = (right + left) / (right - left)
    b = (top + bottom) / (top - bottom)
    c = -far / (far - near)
    d = -far * near / (far - near)
    te = np.zeros(16)
    te[0], te[5], te[8], te[9], te[10], te[11], te[14] = x, y, a, b, c, -1.0, d
    return te


def orthographic(left: float, right: float, top: float, bottom: float, near: float, far: float) -> np.ndarray:
    """Matrix4.makeOrthographic, WebGPU depth range."""
    w = 1.0 / (right - left)
    h = 1.0 / (top - bottom)
    p = 1.0 / (far - near)
    x = (right + left) * w
    y = (top + bottom) * h
    z = near * p
    te = np.zeros(16)
    te[0], te[5], te[10], te[12], te[13], te[14], te[15] = 2 * w, 2 * h, -1 * p, -x, -y, -z, 1.0
    return te


def mat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major a * b."""
    A = np.asarray(a, dtype=np.float64).reshape(4, 4).T
    B = np.asarray(b, dtype=np.float64).reshape(4, 4).T
    return (A @ B).T.reshape(16).copy()


def frustum_planes(proj_view: np.ndarray) -> np.ndarray:
    """Frustum.setFromProjectionMatrix(m, WebGPUCoordinateSystem): 6 x (normal, constant), normalised."""
    me = np.asarray(proj_view, dtype=np.float64)
    rows = [
        (me[3] - me[0], me[7] - me[4], me[11] - me[8], me[15] - me[12]),
        (me[3] + me[0], me[7] + me[4], me[11] + me[8], me[15] + me[12]),
        (me[3] + me[1], me[7] + me[5], me[11] + me[9], me[15] + me[13]),
        (me[3] - me[1], me[7] - me[5], me[11] - me[9], me[15] - me[13]),
        (me[3] - me[2], me[7] - me[6], me[11] - me[10], me[15] - me[14]),
        (me[2], me[6], me[10], me[14]),
    ]
    out = np.zeros((6, 4))
    for i, (x, y, z, w) in enumerate(rows):
        inv = 1.0 / math.sqrt(x * x + y * y + z * z)
        out[i] = (x * inv, y * inv, z * inv, w * inv)
    return out


class Camera:
    """A camera's world matrix, view matrix (matrixWorldInverse) and projection."""

    def __init__(self, projection: np.ndarray, near: float, far: float, perspective_camera: bool = True):
        self.projectionMatrix = np.asarray(projection, dtype=np.float64)
        self.near = near
        self.far = far
        self.isPerspectiveCamera = perspective_camera
        self.matrixWorld = np.eye(4).reshape(16)
        self.matrixWorldInverse = np.eye(4).reshape(16)

    def look_at(self, eye, target, up=(0.0, 1.0, 0.0)):
        self.matrixWorld = look_at_world(eye, target, up)
        self.matrixWorldInverse = invert_rigid(self.matrixWorld)
        return self

    @property
    def position(self) -> np.ndarray:
        return self.matrixWorld[12:15].copy()


def perspective_camera(fov: float, aspect: float, near: float, far: float) -> Camera:
    return Camera(perspective(fov, aspect, near, far), near, far, True)


def orthographic_camera(left, right, top, bottom, near, far) -> Camera:
    return Camera(orthographic(left, right, top, bottom, near, far), near, far, False)


def proj_scale(camera: Camera, viewport_height: float) -> float:
    """projection[1][1] * height / 2 (VirtualGeometry.ts:425)."""
    return float(camera.projectionMatrix[5]) * viewport_height * 0.5
