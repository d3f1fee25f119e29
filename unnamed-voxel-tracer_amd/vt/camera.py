"""Camera block of the render boundary.

``Camera.UniformData`` (src/engine/graphics/camera.zig:12-16) is the 96-byte
std140 block ``u_Camera`` (assets/shaders/camera.glsl:2-6): the camera
position, a column-major ``C_view`` matrix whose columns are the camera's
right, up and forward axes in world space (column 3 = (0,0,0,1)), and the
horizontal field of view in radians.  The reference builds ``C_view`` from
pitch/yaw with zmath; the benchmark workloads use look-at cameras instead,
built here in binary64 and rounded once to binary32.
"""
from __future__ import annotations

import math

import numpy as np

from ._native import Camera as CameraStruct


def look_at(eye, target, fov_deg: float = 70.0, up=(0.0, 1.0, 0.0)) -> CameraStruct:
    eye = np.asarray(eye, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    up = np.asarray(up, dtype=np.float64)
    fwd = target - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(up, fwd)
    n = np.linalg.norm(right)
    if n < 1e-12:  # looking straight up/down: pick any perpendicular
        right = np.cross(np.array([0.0, 0.0, 1.0]), fwd)
        n = np.linalg.norm(right)
    right /= n
    upv = np.cross(fwd, right)
    view = np.zeros(16, dtype=np.float32)
    view[0:3] = right
    view[4:7] = upv
    view[8:11] = fwd
    view[15] = 1.0
    cam = CameraStruct()
    cam.position[:] = [float(np.float32(eye[0])), float(np.float32(eye[1])),
                       float(np.float32(eye[2])), 1.0]
    cam.view[:] = [float(x) for x in view]
    cam.fov = float(np.float32(math.radians(fov_deg)))
    return cam


def camera_bytes(cam: CameraStruct) -> bytes:
    import ctypes
    return ctypes.string_at(ctypes.addressof(cam), ctypes.sizeof(cam))
