"""Camera matrices with the reference's conventions.

* view = inverse of the camera's world transform (Camera.inl:35-40), right-handed, looking down -z;
* projection = glm::perspective(fovy, aspect, near, far) (Camera.inl:42-51), default fovy 0.3*pi;
* view -> uv = fromPlatformNDCToTextureSpace() * projection with the D3D12 texture-space flip
  (RenderAPIAbstractionUtility.h:38-46), uv -> view its inverse (RaytraceCommonResources.cpp:94-99);
* the view-space forward sign is read from the near-plane centre (RaytraceCommonResources.cpp:106-108).
Matrices are computed in float32 as the reference's glm matrices are, and returned row-major
(out = M @ v).
"""
import math

import numpy as np

from . import _lib


def perspective(fovy, aspect, near, far):
    f = 1.0 / math.tan(fovy / 2.0)
    m = np.zeros((4, 4), np.float32)
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -(2.0 * far * near) / (far - near)
    m[3, 2] = -1.0
    return m


def look_at(eye, target, up):
    eye = np.asarray(eye, np.float32)
    f = np.asarray(target, np.float32) - eye
    f = f / np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float32))
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4, dtype=np.float32)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -np.dot(s, eye), -np.dot(u, eye), np.dot(f, eye)
    return m


TO_UV_D3D12 = np.array([[0.5, 0.0, 0.0, 0.5],
                        [0.0, -0.5, 0.0, 0.5],
                        [0.0, 0.0, 1.0, 0.0],
                        [0.0, 0.0, 0.0, 1.0]], np.float32)


class Camera:
    def __init__(self, eye=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fovy=math.pi * 0.3,
                 near=0.1, far=1000.0, aperture_radius=0.0, focus_distance=10.0):
        self.eye, self.target, self.up = eye, target, up
        self.fovy, self.near, self.far = fovy, near, far
        self.aperture_radius, self.focus_distance = aperture_radius, focus_distance

    def matrices(self, width, height):
        view = look_at(self.eye, self.target, self.up)
        proj = perspective(self.fovy, width / float(height), self.near, self.far)
        view_to_uv = (TO_UV_D3D12.astype(np.float64) @ proj.astype(np.float64)).astype(np.float32)
        uv_to_view = np.linalg.inv(view_to_uv.astype(np.float64)).astype(np.float32)
        view_to_world = np.linalg.inv(view.astype(np.float64)).astype(np.float32)
        c = uv_to_view.astype(np.float64) @ np.array([0.5, 0.5, 0.0, 1.0])
        fz = -1.0 if c[2] / c[3] < 0 else 1.0
        return dict(world_to_view=view, view_to_world=view_to_world, view_to_uv=view_to_uv, uv_to_view=uv_to_view,
                    forward_z_sign=fz)


def make_frame(camera, width, height, max_depth=8, env_intensity_scale=0.0, regularize_after=8, flags=0,
               firefly_clamp_scale=12.0, firefly_clamp_relax_power=0.5, kill_nan_samples=1, light_path_vertices=16,
               bdpt_light_nodes=(0, 128), bdpt_camera_nodes=(0, 128)):
    """ypt_frame_desc for a camera; defaults are the reference's renderer variables (RendererVarsList.h:21-47).
    bdpt_light_nodes / bdpt_camera_nodes: the BDPT strategies that contribute (inclusive vertex-count ranges of
    the debug variables _DEBUG.BDPT.EffectiveLightPathNodes / EffectiveCameraPathNodes)."""
    m = camera.matrices(width, height)
    f = _lib.FrameDesc()
    f.width, f.height = width, height
    for name in ("uv_to_view", "view_to_uv", "world_to_view", "view_to_world"):
        getattr(f, name)[:] = [float(v) for v in m[name].ravel()]
    f.aperture_radius = camera.aperture_radius
    f.focus_distance = camera.focus_distance
    f.forward_z_sign = m["forward_z_sign"]
    f.env_intensity_scale = env_intensity_scale
    f.max_depth = max_depth
    f.regularize_after = regularize_after
    f.flags = flags
    f.firefly_clamp_scale = firefly_clamp_scale
    f.firefly_clamp_relax_power = firefly_clamp_relax_power
    f.kill_nan_samples = kill_nan_samples
    f.light_path_vertices = light_path_vertices
    ranges = [int(v) for v in tuple(bdpt_light_nodes) + tuple(bdpt_camera_nodes)]
    f.bdpt_node_ranges[:] = ranges
    if ranges != [0, 128, 0, 128]:
        f.flags |= _lib.YPT_FLAG_BDPT_NODE_RANGES
    return f
