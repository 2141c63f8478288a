"""Integrator-level checks of the CPU checker: closed-form irradiance, white furnace, film merge,
cross-integrator agreement (the IntegratorCompare method) and determinism."""
import math

import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes


def _plane_and_light(albedo=0.5, h=1.0, size=1.0, emission=5.0):
    b = scenes.MeshBuilder()
    fm = scenes.preset_material("PLASTIC")
    fm.albedo[:] = (albedo,) * 3
    fm.specular_amount = 0.0
    fm.roughness = 1.0
    floor = b.material(fm)
    lm = scenes.preset_material("PLASTIC")
    lm.albedo[:] = (0.0, 0.0, 0.0)
    lm.specular_amount = 0.0
    lm.emissive[:] = (emission,) * 3
    light = b.material(lm)
    b.quad((-20, 0, 20), (40, 0, 0), (0, 0, -40), floor)                          # +y
    b.quad((-size / 2, h, -size / 2), (size, 0, 0), (0, 0, size), light)       # -y
    return b.arrays()


def _corner_form_factor(x, y):
    # differential area to a parallel rectangle with one corner above it (x, y = sides / distance)
    a, b = x / math.sqrt(1 + x * x), y / math.sqrt(1 + y * y)
    return (a * math.atan(y / math.sqrt(1 + x * x)) + b * math.atan(x / math.sqrt(1 + y * y))) / (2 * math.pi)


def _mean_y(oracle, scene, cam, integ, spp, w=8, depth=2):
    fr = camera.make_frame(cam, w, w, max_depth=depth, firefly_clamp_scale=0.0)
    r = oracle.render(scene, fr, integ, 0, spp, use_bvh=False)
    return float(r["film"][..., 1].mean())


@pytest.mark.parametrize("integ,spp,tol", [(_lib.YPT_INTEGRATOR_NEE, 256, 0.02), (_lib.YPT_INTEGRATOR_BPT, 1024, 0.03)])
def test_irradiance_under_quad_light(oracle, integ, spp, tol):
    h, size, rho = 1.0, 1.0, 0.5
    scene = _plane_and_light(rho, h, size)
    down = camera.Camera(eye=(0, 0.5, 0), target=(0, 0, 0), up=(0, 0, 1), fovy=0.002)
    up = camera.Camera(eye=(0, 0.5, 0), target=(0, 1, 0), up=(0, 0, 1), fovy=0.002)
    floor_y = _mean_y(oracle, scene, down, integ, spp)
    light_y = _mean_y(oracle, scene, up, integ, spp, depth=1)
    F = 4 * _corner_form_factor(size / 2 / h, size / 2 / h)
    assert abs(floor_y / light_y - rho * F) / (rho * F) < tol, (floor_y / light_y, rho * F)


def _furnace_scene(kind):
    b = scenes.MeshBuilder()
    if kind == "diffuse":
        m = scenes.preset_material("PLASTIC")
        m.albedo[:] = (1.0, 1.0, 1.0)
        m.specular_amount = 0.0
        m.roughness = 1.0
    elif kind == "conductor":
        m = scenes.preset_material("METAL_SILVER")
        m.albedo[:] = (1.0, 1.0, 1.0)
        m.specular[:] = (1.0, 1.0, 1.0)
        m.roughness = 0.6
    elif kind == "plastic":
        m = scenes.preset_material("PLASTIC")
        m.albedo[:] = (1.0, 1.0, 1.0)
        m.roughness = 0.5
    else:
        m = scenes.preset_material("GLASS")
        m.roughness = 0.4
    mat = b.material(m)
    P, N, I = scenes.sphere_template(24, 24)
    b.add(P, N, I, mat)
    return b.arrays()


@pytest.mark.parametrize("kind,lo,hi", [("diffuse", 0.99, 1.01), ("conductor", 0.90, 1.01), ("plastic", 0.93, 1.01),
                                        ("glass", 0.90, 1.05)])
def test_white_furnace(oracle, kind, lo, hi):
    # WHITE_FURNACE_TEST semantics: the environment and every other object emit unit white light
    scene = _furnace_scene(kind)
    cam = camera.Camera(eye=(0, 0, 3.2), target=(0, 0, 0), fovy=0.75)
    fr = camera.make_frame(cam, 48, 48, max_depth=64, flags=_lib.YPT_FLAG_WHITE_FURNACE, firefly_clamp_scale=0.0)
    r = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 0, 64, use_bvh=False)
    y = r["film"][..., 1]
    yy, xx = np.mgrid[0:48, 0:48]
    rr = np.hypot(xx - 23.5, yy - 23.5)
    sphere, back = y[rr < 12], y[rr > 22]
    ratio = sphere.mean() / back.mean()
    assert lo <= ratio <= hi, ratio


def test_film_is_running_mean_of_passes(oracle, cornell):
    scene, cam = cornell
    fr = camera.make_frame(cam, 24, 24, max_depth=8, firefly_clamp_scale=0.0)
    passes = [oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, k, 1, use_bvh=False)["last"][..., :3] for k in range(6)]
    film = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 0, 6, use_bvh=False)
    assert np.allclose(film["film"][..., :3], np.mean(passes, axis=0), rtol=1e-5, atol=1e-7)
    assert np.all(film["counts"] == 6)


def test_nan_samples_are_dropped_and_clamp_limits_fireflies(oracle, cornell):
    # the clamp only ever scales a sample down, toward the 3x3 neighbourhood of the previous film
    scene, cam = cornell
    fr = camera.make_frame(cam, 24, 24, max_depth=8, firefly_clamp_scale=1.0, firefly_clamp_relax_power=0.0)
    clamped = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 0, 8, use_bvh=False)["film"]
    fr.firefly_clamp_scale = 0.0
    free = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 0, 8, use_bvh=False)["film"]
    assert np.all(clamped[..., :3] <= free[..., :3] + 1e-6)
    assert clamped[..., :3].sum() < free[..., :3].sum()


def _rel_d_16x(a, b):
    # IntegratorCompare rel|d| 16x: mean |a - b| after box-downsampling, relative to the mean brightness
    h, w = a.shape[:2]
    f = 4
    A = a[:h - h % f, :w - w % f].reshape(h // f, f, w // f, f, -1).mean((1, 3))
    B = b[:h - h % f, :w - w % f].reshape(h // f, f, w // f, f, -1).mean((1, 3))
    return float(np.abs(A - B).mean() / (0.5 * (A.mean() + B.mean())))


def test_bpt_and_nee_converge_to_the_same_image(oracle, cornell):
    scene, cam = cornell
    fr = camera.make_frame(cam, 48, 48, max_depth=8, firefly_clamp_scale=0.0)
    nee = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 0, 256)["film"][..., :3]
    bpt = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_BPT, 0, 256)["film"][..., :3]
    assert abs(nee.mean() - bpt.mean()) / nee.mean() < 0.02
    assert _rel_d_16x(nee, bpt) < 0.08


def test_render_is_independent_of_thread_count(oracle, cornell):
    scene, cam = cornell
    fr = camera.make_frame(cam, 40, 40, max_depth=8)
    a = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 3, 3, threads=1)["film"]
    b = oracle.render(scene, fr, _lib.YPT_INTEGRATOR_NEE, 3, 3, threads=8)["film"]
    assert np.array_equal(a, b)


def test_oracle_bvh_matches_brute_force(oracle, shapes_small):
    scene = shapes_small[0]
    rng = np.random.default_rng(9)
    n = 3000
    o = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 3], rays[:, 4:7], rays[:, 7] = o, 5e-4, d, np.inf
    t0, i0, _ = oracle.trace_closest(scene, rays, use_bvh=False)
    t1, i1, _ = oracle.trace_closest(scene, rays, use_bvh=True)
    assert np.array_equal(i0, i1) and np.array_equal(t0.view(np.uint32), t1.view(np.uint32))
    assert (i0 != 0xFFFFFFFF).mean() > 0.2


def test_watertight_shared_edges(oracle):
    # rays through the shared edge and vertex of a fan of triangles never slip through
    b = scenes.MeshBuilder()
    m = b.material(scenes.preset_material("PLASTIC"))
    k = 12
    ang = np.linspace(0, 2 * np.pi, k + 1)[:-1]
    P = [[0, 0, 0]] + [[np.cos(a), 0, np.sin(a)] for a in ang]
    I = [[0, 1 + (i + 1) % k, 1 + i] for i in range(k)]
    b.add(np.array(P, np.float32), np.tile([0, 1, 0], (k + 1, 1)), np.array(I), m)
    scene = b.arrays()
    rays = []
    for a in np.concatenate([ang, ang + 1e-7]):
        for r in (0.0, 0.25, 0.5):
            x, z = r * np.cos(a), r * np.sin(a)
            rays.append([x, 1.0, z, 5e-4, 0, -1, 0, np.inf])
    t, tri, _ = oracle.trace_closest(scene, np.array(rays, np.float32), use_bvh=False)
    assert np.all(tri != 0xFFFFFFFF)
    assert np.allclose(t, 1.0, atol=1e-6)
