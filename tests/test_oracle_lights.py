"""Punctual lights (point, spot, directional) and environment light subpaths on the CPU checker:
closed-form irradiance of each light type on a diffuse plane, BPT equal to NEE bit for bit when only
punctual lights exist (both sample them from the same alias table), and the bidirectional integrator
converging to the NEE image on the punctual-light test scene
(Tests/IntegratorCompare/punctual_lights_scene.py; the reference reports 0.1-0.8% agreement)."""
import math

import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes

NEE, BPT, BDPT = _lib.YPT_INTEGRATOR_NEE, _lib.YPT_INTEGRATOR_BPT, _lib.YPT_INTEGRATOR_BDPT
RHO = 0.5


def _plane(lights, emitter=False):
    b = scenes.MeshBuilder()
    fm = scenes.preset_material("PLASTIC")
    fm.albedo[:] = (RHO,) * 3
    fm.specular_amount = 0.0
    fm.roughness = 1.0
    b.quad((-50, 0, 50), (100, 0, 0), (0, 0, -100), b.material(fm))   # +y
    if emitter:   # grey emitter of radiance 1 far below the floor, facing down: calibrates Y of unit radiance
        lm = scenes.blank_material()
        lm.emissive[:] = (1.0, 1.0, 1.0)
        b.quad((-1, -10, -1), (0, 0, 2), (2, 0, 0), b.material(lm))
    return b.arrays(lights)


def _light(t, pos, target, intensity, inner=0.0, outer=math.pi * 0.25):
    l = _lib.Light(t)
    l.lookAt(pos, target)
    l.setColor((1.0, 1.0, 1.0))
    l.setIntensity(intensity)
    l.setSpotConeAngles(inner, outer)
    return l


def _y_at(oracle, scene, x, z, integ=NEE, spp=4, eye_y=0.5, depth=1):
    cam = camera.Camera(eye=(x, eye_y, z), target=(x, eye_y - 1.0, z), up=(0, 0, 1), fovy=0.002)
    fr = camera.make_frame(cam, 4, 4, max_depth=depth, firefly_clamp_scale=0.0)
    return float(oracle.render(scene, fr, integ, 0, spp, use_bvh=False)["film"][..., 1].mean())


@pytest.fixture(scope="module")
def unit_radiance_y(oracle):
    scene = _plane([], emitter=True)
    cam = camera.Camera(eye=(0, -9, 0), target=(0, -10, 0), up=(0, 0, 1), fovy=0.002)
    fr = camera.make_frame(cam, 4, 4, max_depth=1, firefly_clamp_scale=0.0)
    return float(oracle.render(scene, fr, NEE, 0, 4, use_bvh=False)["film"][..., 1].mean())


@pytest.mark.parametrize("integ", [NEE, BPT])
def test_point_light_inverse_square_and_cosine(oracle, unit_radiance_y, integ):
    h, I = 2.0, 40.0
    scene = _plane([_light(_lib.Light.POINT, (0, h, 0), (0, 0, 0), I)])
    for r in (0.0, h, 2.5 * h):
        want = RHO / math.pi * I * h / (h * h + r * r) ** 1.5   # E = I cos / d^2, L = rho / pi E
        got = _y_at(oracle, scene, r, 0.0, integ) / unit_radiance_y
        assert abs(got / want - 1) < 2e-3, (r, got, want)


def test_spot_light_cone_falloff(oracle, unit_radiance_y):
    h, I, inner, outer = 2.0, 60.0, 0.3, 0.5
    scene = _plane([_light(_lib.Light.SPOT, (0, h, 0), (0, 0, 0), I, inner, outer)])
    ci, co = math.cos(inner), math.cos(outer)
    for ang in (0.0, 0.25, 0.35, 0.4, 0.45, 0.6):
        r = h * math.tan(ang)
        c = math.cos(ang)
        t = min(max((c - co) / (ci - co), 0.0), 1.0)
        falloff = t * t * (3 - 2 * t)                     # HLSL smoothstep(cosOuter, cosInner, cos)
        want = RHO / math.pi * I * falloff * h / (h * h + r * r) ** 1.5
        got = _y_at(oracle, scene, r, 0.0) / unit_radiance_y
        assert abs(got - want) <= 2e-3 * max(want, 1e-3), (ang, got, want)


def test_directional_light_cosine(oracle, unit_radiance_y):
    I = 3.0
    for theta in (0.0, 0.7):
        d = (math.sin(theta), -math.cos(theta), 0.0)
        scene = _plane([_light(_lib.Light.DIRECTIONAL, (0, 5, 0), (d[0], 5 + d[1], d[2]), I)])
        want = RHO / math.pi * I * math.cos(theta)
        got = _y_at(oracle, scene, 1.0, -2.0) / unit_radiance_y
        assert abs(got / want - 1) < 2e-3, (theta, got, want)


def test_occluded_punctual_light_adds_nothing(oracle):
    b = scenes.MeshBuilder()
    fm = scenes.preset_material("PLASTIC")
    fm.albedo[:] = (RHO,) * 3
    fm.specular_amount = 0.0
    fm.roughness = 1.0
    mat = b.material(fm)
    b.quad((-50, 0, 50), (100, 0, 0), (0, 0, -100), mat)
    b.quad((-1, 1, -1), (2, 0, 0), (0, 0, 2), mat)     # a blocker at y = 1, facing down
    scene = b.arrays([_light(_lib.Light.POINT, (0, 2, 0), (0, 0, 0), 40.0)])
    assert _y_at(oracle, scene, 0.0, 0.0, eye_y=0.5) == 0.0
    assert _y_at(oracle, scene, 3.0, 0.0, eye_y=0.5) > 0.0


@pytest.mark.parametrize("light_type", ["point", "spot", "directional"])
def test_bpt_equals_nee_with_punctual_lights_only(oracle, light_type):
    # with no mesh lights both integrators pick among the same punctual lights with the same table
    scene, cam, env = scenes.punctual_lights(light_type, sphere_res=(12, 12))
    fr = camera.make_frame(cam, 32, 20, max_depth=6, env_intensity_scale=env)
    a = oracle.render(scene, fr, BPT, 0, 4)["film"]
    b = oracle.render(scene, fr, NEE, 0, 4)["film"]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert a[..., 1].max() > 0


@pytest.mark.parametrize("light_type", ["point", "spot", "directional", "mixed", "env"])
def test_bdpt_converges_to_nee_with_punctual_and_environment_lights(oracle, light_type):
    scene, cam, env = scenes.punctual_lights(light_type, sphere_res=(12, 12))
    # 64x40 gives 8x5 light subpaths per pass: fewer make the light-tracing strategies too noisy
    fr = camera.make_frame(cam, 64, 40, max_depth=8, env_intensity_scale=env, firefly_clamp_scale=0.0)
    nee = oracle.render(scene, fr, NEE, 0, 256)["film"][..., :3].astype(np.float64)
    bd = oracle.render(scene, fr, BDPT, 0, 256)["film"][..., :3].astype(np.float64)
    ratio = bd.mean() / nee.mean()
    assert abs(ratio - 1) < 0.01, ratio
