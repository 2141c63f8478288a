"""C ABI surface and host-side logic (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "yapt_pt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ypt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(ylib):
    lib = ylib.load()
    names = _declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(ylib.EXPORTED_SYMBOLS) | {"ypt_version"}


def test_host_sobol_point_matches_checker(ylib, oracle):
    for idx in (0, 5, 1 << 20):
        assert np.array_equal(ylib.sobol_point(idx), oracle.sobol_point(idx))


def test_errors_are_codes_not_crashes(ylib):
    lib = ylib.load()
    h = ctypes.c_void_p()
    assert lib.ypt_create(7, 0, ctypes.byref(h)) == -1
    assert b"unknown integrator" in lib.ypt_last_error()
    assert lib.ypt_create(2, 0, ctypes.byref(h)) < 0        # BDPT handle needs a device, or is refused
    assert lib.ypt_render(None, 0, 1, None) == -1
    assert lib.ypt_set_option(None, b"bvh_width", 2) == -1
    assert lib.ypt_destroy(None) == 0


def test_version_string(ylib):
    assert b"gfx950" in ylib.load().ypt_version()


def test_cornell_geometry_conventions(cornell):
    scene, cam = cornell
    assert scene.triangle_count == 36
    p, i = scene.positions, scene.indices
    n = np.cross(p[i[:, 1]] - p[i[:, 0]], p[i[:, 2]] - p[i[:, 0]])
    centre = np.array([0.0, 1.0, 0.0], np.float32)
    c = p[i].mean(axis=1)
    walls = slice(0, 10)
    # walls face the room interior: front face = dot(ray, cross(p2 - p1, p3 - p1)) < 0 from inside
    assert np.all(np.einsum("ij,ij->i", n[walls], centre - c[walls]) > 0)
    # the light faces down
    assert np.all(n[-2:, 1] < 0)
    assert np.allclose(np.linalg.norm(scene.tangents[:, :3], axis=1), 1.0, atol=1e-6)
    assert np.allclose(np.einsum("ij,ij->i", scene.tangents[:, :3], scene.normals), 0.0, atol=1e-6)


def test_random_shapes_size():
    from yapt_amd import scenes
    scene, _ = scenes.random_shapes()
    assert 0.95e6 < scene.triangle_count < 1.1e6
    assert len(scene.materials) == 2002
    glass = sum(1 for m in scene.materials if m.transparency == 1.0)
    assert 150 < glass < 250


def test_camera_matrices_round_trip():
    from yapt_amd import camera
    cam = camera.Camera(eye=(1.0, 2.0, 3.0), target=(0.0, 0.5, 0.0), fovy=math.pi * 0.3)
    m = cam.matrices(320, 200)
    assert np.allclose(m["uv_to_view"] @ m["view_to_uv"], np.eye(4), atol=1e-5)
    assert np.allclose(m["view_to_world"] @ m["world_to_view"], np.eye(4), atol=1e-5)
    assert m["forward_z_sign"] == -1.0
    # the centre of the film looks at the target, the top edge is half the vertical field of view up
    def ray(u, v):
        p = m["uv_to_view"] @ np.array([u, v, 0.0, 1.0])
        d = p[:3] / p[3]
        d = m["view_to_world"][:3, :3] @ d
        return d / np.linalg.norm(d)
    fwd = np.array([-1.0, -1.5, -3.0]) / np.linalg.norm([-1.0, -1.5, -3.0])
    assert np.allclose(ray(0.5, 0.5), fwd, atol=1e-5)
    assert abs(math.acos(np.dot(ray(0.5, 0.0), fwd)) - math.pi * 0.15) < 1e-4
    assert ray(0.5, 0.0)[1] > fwd[1]   # v = 0 is the top of the image (D3D texture space)


def test_denoiser_resampling_rejects_odd_and_empty_shapes():
    """Pooling and upsampling entry points validate the full-resolution shape on the host: an odd H or
    W would index past the half-resolution tensor. The fake pointers are never dereferenced, since the
    checks return before any launch."""
    from yapt_amd import denoiser
    lib = denoiser.load()
    p = ctypes.c_void_p(256)
    for H, W in ((7, 8), (8, 7), (0, 8)):
        assert lib.ydn_pool2(p, 1, H, W, 16, p, None) == -1
        assert lib.ydn_pool2_grad(p, 1, H, W, 16, p, None) == -1
        assert lib.ydn_up2_concat(p, 16, p, 16, 1, H, W, p, None) == -1
        assert lib.ydn_up2_concat_grad(p, 16, 16, 1, H, W, p, p, 0, None) == -1
    assert b"even H and W" in lib.ydn_last_error() or b"empty shape" in lib.ydn_last_error()
    assert lib.ydn_up2_concat(None, 16, p, 16, 1, 8, 8, p, None) == -1
    assert b"null tensor" in lib.ydn_last_error()
    for H, W, c1, c2, cout in ((7, 8, 32, 32, 32), (8, 8, 24, 32, 32), (8, 8, 32, 32, 48), (0, 8, 32, 32, 32)):
        assert lib.ydn_conv3x3_up2cat(p, c1, p, c2, 1, H, W, p, None, cout, 1, p, None) == -1
    assert lib.ydn_conv3x3(p, 1, 8, 8, 24, p, None, 32, 1, p, None) == -1
    assert b"multiple of 16" in lib.ydn_last_error()
    # the bias of one conv is staged in LDS, 256 channels at most
    assert lib.ydn_conv3x3(p, 1, 8, 8, 32, p, p, 288, 1, p, None) == -1
    assert b"at most 256" in lib.ydn_last_error()
    assert lib.ydn_conv3x3_pool2(p, 1, 8, 8, 32, p, p, 288, 1, p, p, None) == -1
    assert lib.ydn_conv3x3_up2cat(p, 16, p, 16, 1, 8, 8, p, p, 288, 1, p, None) == -1


def test_denoiser_library_exports_every_declared_function():
    import re
    from yapt_amd import denoiser
    lib = denoiser.load()
    with open(os.path.join(ROOT, "include", "yapt_denoise.h")) as f:
        names = re.findall(r"\b(ydn_\w+)\(", f.read())
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
