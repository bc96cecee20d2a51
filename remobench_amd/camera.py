"""Camera poses and SharedUniforms for the per-frame passes.

Follows the reference shell's camera (src/shell/OrbitControls.h:68-93,
src/shell/GLRenderer.cpp:74-78, src/shell/App.cpp:200-248): fovy 60 degrees,
near = extent * 1e-4, far = extent * 100, a GLM right-handed perspective with
[-1, 1] depth, transform = proj * view, handed to the device transposed into
row-major rows[r][c] (toDeviceMat, src/shell/App.cpp:23-31). Matrices are
built in float64 and rounded to float32 once, as the reference does.

Poses on the seed-generated scenes:
  overview -- OrbitControls::frameBox: target the box centre, yaw 0.35,
              pitch -0.6, radius = (|box| / 2) / tan(fovy / 2) * 1.1;
  street   -- eye 1.7 m above the ground, 10 m from a ground target near the
              scene centre;
  closeup  -- 2 m in front of the cloud's point nearest the scene centre at
              1-2 m height (a box face or a pole).
"""
from __future__ import annotations

import math

import numpy as np

from . import lib as L

FOVY_DEG = 60.0


def perspective(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -(2.0 * far * near) / (far - near)
    m[3, 2] = -1.0
    return m


def translate(v) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = v
    return m


def rotate(angle: float, axis) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = math.cos(angle), math.sin(angle)
    x, y, z = a
    r = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    m = np.eye(4)
    m[:3, :3] = r
    return m


def orbit_world(target, yaw: float, pitch: float, radius: float) -> np.ndarray:
    """OrbitControls::update: world = T(target) R(yaw, z) R(pitch, x) flip T(0, 0, r)."""
    flip = np.array([[1, 0, 0, 0], [0, 0, -1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.float64)
    return (translate(target) @ rotate(yaw, (0, 0, 1)) @ rotate(pitch, (1, 0, 0)) @ flip
            @ translate((0, 0, radius)))


def look_at(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """glm::lookAt (right-handed): the view matrix."""
    eye, target, up = (np.asarray(v, dtype=np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def _mat(dst: L.Mat4, m: np.ndarray):
    m32 = m.astype(np.float32)
    for r in range(4):
        for c in range(4):
            dst.rows[r][c] = float(m32[r, c])


def uniforms(view: np.ndarray, box_size, width: int = 1920, height: int = 1080,
             lod_pixel_budget: float = 128.0, point_size: int = 1) -> L.SharedUniforms:
    """SharedUniforms as App::buildUniforms fills them (src/shell/App.cpp:215-248)."""
    extent = max(max(float(v) for v in box_size), 1.0)
    near, far = extent * 1e-4, extent * 100.0
    proj = perspective(math.radians(FOVY_DEG), width / height, near, far)
    transform = proj @ view
    u = L.SharedUniforms()
    u.width, u.height = float(width), float(height)
    u.fovy_rad = float(math.radians(FOVY_DEG))
    _mat(u.view, view)
    _mat(u.proj, proj)
    _mat(u.transform, transform)
    _mat(u.transform_frozen, transform)
    _mat(u.transform_frozen_inv, np.linalg.inv(transform))
    u.box_min.x = u.box_min.y = u.box_min.z = 0.0
    u.box_max.x, u.box_max.y, u.box_max.z = (float(v) for v in box_size)
    u.lod_pixel_budget = lod_pixel_budget
    u.min_node_size = 64.0
    u.lod_scale = 0.5
    u.point_size = point_size
    u.color_mode = 0
    u.edl_strength = 0.4
    u.do_update_visibility = 1
    u.enable_edl = 1
    u.show_points = 1
    return u


def overview_view(box_size) -> np.ndarray:
    bmax = np.array([float(v) for v in box_size])
    target = bmax * 0.5
    extent = np.linalg.norm(bmax)
    radius = (extent * 0.5) / math.tan(max(math.radians(FOVY_DEG), 1e-3) * 0.5) * 1.1
    return np.linalg.inv(orbit_world(target, 0.35, -0.6, radius))


def street_view(box_size, ground_z: float) -> np.ndarray:
    c = np.array([float(box_size[0]) * 0.5, float(box_size[1]) * 0.5, ground_z])
    eye = c + np.array([-10.0 / math.sqrt(2), -10.0 / math.sqrt(2), 1.7])
    return look_at(eye, c)


def closeup_view(points: np.ndarray, box_size) -> np.ndarray:
    c = np.array([float(box_size[0]) * 0.5, float(box_size[1]) * 0.5])
    z = points["z"].astype(np.float64)
    zmin = z.min()
    sel = np.nonzero((z > zmin + 1.0) & (z < zmin + 2.0))[0]
    if len(sel) == 0:
        sel = np.arange(len(points))
    d = (points["x"][sel] - c[0]) ** 2 + (points["y"][sel] - c[1]) ** 2
    p = points[sel[int(np.argmin(d))]]
    target = np.array([p["x"], p["y"], p["z"]], dtype=np.float64)
    horiz = target[:2] - c
    n = np.linalg.norm(horiz)
    direction = horiz / n if n > 1e-6 else np.array([1.0, 0.0])
    eye = target + np.array([direction[0] * 2.0, direction[1] * 2.0, 0.3])
    return look_at(eye, target)


def poses(points: np.ndarray, box_size, width: int = 1920, height: int = 1080) -> dict:
    ground = float(np.median(points["z"][: min(len(points), 200_000)]))
    return {
        "overview": uniforms(overview_view(box_size), box_size, width, height),
        "street": uniforms(street_view(box_size, ground), box_size, width, height),
        "closeup": uniforms(closeup_view(points, box_size), box_size, width, height),
    }
