"""Host-side 4x4 matrix helpers, column-major like WGSL mat4x4<f32>.

They follow the formulas of the reference's matrix library (wgpu-matrix 2.9.0,
used at src/utils/matrices.ts:11-48 and camera.ts:61-110): WebGPU clip space
with z in [0, 1]. Values are computed in float64 and stored as float32, as the
library stores into Float32Array. The kernels never rebuild matrices; they
read the packed uniform bytes, so the oracle and the HIP path see the same
values.
"""
from __future__ import annotations

import math

import numpy as np


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float64)


def to_f32_colmajor(m: np.ndarray) -> np.ndarray:
    """Row/column-indexed 4x4 (m[row, col]) -> 16 float32 in column-major order."""
    return np.ascontiguousarray(m.T.reshape(16)).astype(np.float32)


def perspective(fovy_rad: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = math.tan(math.pi * 0.5 - 0.5 * fovy_rad)
    range_inv = 1.0 / (near - far)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = far * range_inv
    m[3, 2] = -1.0
    m[2, 3] = far * near * range_inv
    return m


def ortho(left, right, bottom, top, near, far) -> np.ndarray:
    m = np.zeros((4, 4))
    m[0, 0] = 2.0 / (right - left)
    m[1, 1] = 2.0 / (top - bottom)
    m[2, 2] = 1.0 / (near - far)
    m[0, 3] = (right + left) / (left - right)
    m[1, 3] = (top + bottom) / (bottom - top)
    m[2, 3] = near / (near - far)
    m[3, 3] = 1.0
    return m


def _norm(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def look_at(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """View matrix (world -> camera), as wgpu-matrix mat4.lookAt."""
    eye = np.asarray(eye, dtype=np.float64)
    z = _norm(eye - np.asarray(target, dtype=np.float64))
    x = _norm(np.cross(up, z))
    y = _norm(np.cross(z, x))
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = x, y, z
    m[0, 3], m[1, 3], m[2, 3] = -x @ eye, -y @ eye, -z @ eye
    return m


def translation(t) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = t[:3]
    return m


def scale(m: np.ndarray, s) -> np.ndarray:
    return m @ np.diag([s[0], s[1], s[2], 1.0])


def invert(m: np.ndarray) -> np.ndarray:
    return np.linalg.inv(m)


def spherical_to_cartesian(phi_deg: float, theta_deg: float) -> np.ndarray:
    """src/utils/index.ts:60-74"""
    phi, theta = math.radians(phi_deg), math.radians(theta_deg)
    return np.array([math.cos(phi) * math.sin(theta), math.cos(theta), math.sin(phi) * math.sin(theta)])


def project_point(m: np.ndarray, p) -> np.ndarray:
    v = np.array([p[0], p[1], p[2], 1.0])
    return m @ v
