"""Camera math in three.js conventions (fp64, column-major 16-element matrices).

VirtualGeometry.update() (src/core/runtime/VirtualGeometry.ts:414-437) reads the camera's
matrixWorldInverse and projection matrix, derives projScale = P[5] * height / 2 and extracts the
frustum planes with Frustum.setFromProjectionMatrix(P * V, WebGPUCoordinateSystem). These are
restated here from three.js's published formulas (PerspectiveCamera / OrthographicCamera
updateProjectionMatrix, Matrix4.makePerspective / makeOrthographic with the WebGPU depth range [0, 1],
Matrix4.lookAt, Matrix4.multiplyMatrices, Frustum.setFromProjectionMatrix), in their operation order
and with scalar arithmetic (no library dot products, whose summation order is unspecified), so the
Node host's restatement (js/virtual_geometry.js) gives the same doubles. The camera's rotation is kept
as the lookAt basis and its view matrix is the rigid inverse (three.js goes through a quaternion and a
general 4x4 inverse, which can differ in the last bits); three.js itself is absent, so bit-equality with
it is not claimed. The values are pinned by closed forms (tests/test_oracle_pins.py).
"""
from __future__ import annotations

import math

import numpy as np

DEG2RAD = math.pi / 180


def _normalize3(x: float, y: float, z: float):
    """Vector3.normalize: divideScalar(length || 1) = multiplyScalar(1 / length)."""
    n = math.sqrt(x * x + y * y + z * z)
    inv = 1.0 / (n or 1.0)
    return x * inv, y * inv, z * inv


def _cross(ax, ay, az, bx, by, bz):
    """Vector3.crossVectors(a, b)."""
    return ay * bz - az * by, az * bx - ax * bz, ax * by - ay * bx


def look_at_world(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """World matrix of a camera at `eye` looking at `target` (its -z axis points at the target): the rotation
    Matrix4.lookAt(eye, target, up) builds, and the translation `eye`."""
    ex, ey, ez = (float(v) for v in eye)
    tx, ty, tz = (float(v) for v in target)
    ux, uy, uz = (float(v) for v in up)
    zx, zy, zz = ex - tx, ey - ty, ez - tz
    if zx * zx + zy * zy + zz * zz == 0:
        zz = 1.0
    zx, zy, zz = _normalize3(zx, zy, zz)
    xx, xy, xz = _cross(ux, uy, uz, zx, zy, zz)
    if xx * xx + xy * xy + xz * xz == 0:
        # up parallel to the view direction: nudged as three.js does
        if abs(uz) == 1:
            zx += 0.0001
        else:
            zz += 0.0001
        zx, zy, zz = _normalize3(zx, zy, zz)
        xx, xy, xz = _cross(ux, uy, uz, zx, zy, zz)
    xx, xy, xz = _normalize3(xx, xy, xz)
    yx, yy, yz = _cross(zx, zy, zz, xx, xy, xz)
    return np.array([xx, xy, xz, 0.0, yx, yy, yz, 0.0, zx, zy, zz, 0.0, ex, ey, ez, 1.0])


def invert_rigid(world: np.ndarray) -> np.ndarray:
    """Inverse of a rotation + translation (column-major): R^T and -R^T t, sums left to right."""
    m = [float(v) for v in np.asarray(world, dtype=np.float64).reshape(16)]
    out = [0.0] * 16
    for i in range(3):  # row i of R^T = column i of R
        out[0 + i] = m[i * 4 + 0]
        out[4 + i] = m[i * 4 + 1]
        out[8 + i] = m[i * 4 + 2]
        out[12 + i] = -((m[i * 4 + 0] * m[12] + m[i * 4 + 1] * m[13]) + m[i * 4 + 2] * m[14])
    out[15] = 1.0
    return np.array(out)


def perspective(fov_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    """PerspectiveCamera.updateProjectionMatrix (zoom 1, no view offset) + makePerspective, WebGPU depth range."""
    top = near * math.tan(DEG2RAD * 0.5 * fov_deg)
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
    """OrthographicCamera.updateProjectionMatrix (zoom 1, no view offset) + makeOrthographic, WebGPU depth range."""
    dx = (right - left) / 2
    dy = (top - bottom) / 2
    cx = (right + left) / 2
    cy = (top + bottom) / 2
    left, right, top, bottom = cx - dx, cx + dx, cy + dy, cy - dy
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
    """Column-major a * b, as Matrix4.multiplyMatrices sums: a_i1 b_1j + a_i2 b_2j + a_i3 b_3j + a_i4 b_4j."""
    A = [float(v) for v in np.asarray(a, dtype=np.float64).reshape(16)]
    B = [float(v) for v in np.asarray(b, dtype=np.float64).reshape(16)]
    out = [0.0] * 16
    for j in range(4):
        for i in range(4):
            out[j * 4 + i] = A[i] * B[j * 4] + A[4 + i] * B[j * 4 + 1] + A[8 + i] * B[j * 4 + 2] + A[12 + i] * B[j * 4 + 3]
    return np.array(out)


def frustum_planes(proj_view: np.ndarray) -> np.ndarray:
    """Frustum.setFromProjectionMatrix(m, WebGPUCoordinateSystem): 6 x (normal, constant), normalised
    (Plane.normalize: everything times 1 / |normal|)."""
    me = [float(v) for v in np.asarray(proj_view, dtype=np.float64).reshape(16)]
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
