"""Subsurface scattering (yapt_amd/csrc/core/subsurface.h) on the checker: the Burley profile against its
normalisation and the closed-form CDF of its radius sampler, the searchlight fit of d, the exit-point
density's axis and lane mixture on closed-form configurations, energy under the white furnace (a
subsurface plane returns what a Lambertian plane of the same albedo returns, less the profile's tail
beyond r_max), BPT and BDPT against NEE on a lit subsurface scene, and on the GPU the subsurface
kernels against the checker for all three integrators (both forms of the BDPT camera stage)."""
import math

import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes

NEE, BPT, BDPT = _lib.YPT_INTEGRATOR_NEE, _lib.YPT_INTEGRATOR_BPT, _lib.YPT_INTEGRATOR_BDPT
RMAX_FACTOR = 3 * math.log(1000)


def _mixture_cdf(r, d):
    """radius CDF of the profile: (1/4) Exp(mean d) + (3/4) Exp(mean 3d)"""
    return 1 - 0.25 * np.exp(-r / d) - 0.75 * np.exp(-r / (3 * d))


def test_profile_is_normalised_over_the_plane(oracle):
    d = 0.3
    r = np.concatenate([np.geomspace(1e-6, 0.1, 4000), np.linspace(0.1, 40 * d, 40000)[1:]])
    R = oracle.evaluate("burley_profile_area", np.stack([r, np.full_like(r, d)], 1))[:, 0].astype(np.float64)
    f = 2 * np.pi * r * R
    total = np.trapezoid(f, r) + 2 * np.pi * 1e-6 * (1 / (4 * np.pi * d)) * 1e-6 / 2   # the first sliver
    assert abs(total - 1.0) < 2e-4, total
    inside = r < RMAX_FACTOR * d
    mass = np.trapezoid(f[inside], r[inside])
    assert abs(mass - float(_mixture_cdf(RMAX_FACTOR * d, d))) < 2e-4      # 0.99925 within r_max


def test_radius_sampler_inverts_the_mixture_cdf(oracle):
    d = 0.7
    u = (np.arange(20000) + 0.5) / 20000
    r = oracle.evaluate("burley_sample_radius", np.stack([u, np.full_like(u, d)], 1))[:, 0].astype(np.float64)
    F = _mixture_cdf(np.sort(r), d)
    assert np.max(np.abs(F - np.sort(u))) < 2e-3                           # Kolmogorov distance


def test_searchlight_fit_of_d(oracle):
    A = np.array([0.0, 0.3, 0.8, 1.0])
    rows = np.stack([np.full(4, 0.5), A], 1)
    d = oracle.evaluate("burley_d", rows)[:, 0]
    s = 1.85 - A + 7 * np.abs(A - 0.8) ** 3                                # Christensen & Burley 2015
    assert np.allclose(d, 0.5 / s, rtol=1e-6)


def test_exit_density_mixes_axes_and_lanes(oracle):
    albedo, radius = 0.5, 0.2
    d = 0.2 / (1.85 - 0.5 + 7 * 0.3 ** 3)
    Rrow = lambda r: oracle.evaluate("burley_profile_area", [[r, d]])[0, 0]
    # exit on the entry plane (normal +y): only the normal axis (probability 1/2) reaches it
    x, z = 0.03, -0.04
    p = oracle.evaluate("pdf_subsurface_exit_plane", [[albedo, radius, x, 0, z, 0, 1, 0]])[0, 0]
    assert abs(p - 0.5 * Rrow(math.hypot(x, z))) < 1e-6 * p
    # an exit facing +z: only the tangent axis of the normal's basis (probability 1/4) reaches it, the
    # disk of that axis spans x and y
    p = oracle.evaluate("pdf_subsurface_exit_plane", [[albedo, radius, x, 0.02, z, 0, 0, 1]])[0, 0]
    assert abs(p - 0.25 * Rrow(math.hypot(x, 0.02))) < 1e-6 * p
    # beyond r_max of every lane the density is zero
    far = RMAX_FACTOR * d * 1.01
    assert oracle.evaluate("pdf_subsurface_exit_plane", [[albedo, radius, far, 0, 0, 0, 1, 0]])[0, 0] == 0.0


def _plane(material, size=20.0):
    b = scenes.MeshBuilder()
    b.quad((-size, 0, size), (2 * size, 0, 0), (0, 0, -2 * size), b.material(material))
    return b


def _sss_material(albedo=0.6, radius=0.05, amount=1.0):
    m = scenes.preset_material("PLASTIC")
    m.albedo[:] = (albedo,) * 3
    m.specular_amount = 0.0
    m.roughness = 1.0
    m.subsurface_amount = amount
    m.subsurface_radius[:] = (radius,) * 3
    m.subsurface_radius_scale = 1.0
    return m


def test_subsurface_plane_conserves_energy_in_the_furnace(oracle):
    cam = camera.Camera(eye=(0, 3, 0), target=(0, 0, 0), up=(0, 0, -1), fovy=0.6)
    fr = camera.make_frame(cam, 48, 48, max_depth=3, firefly_clamp_scale=0.0, flags=_lib.YPT_FLAG_WHITE_FURNACE)
    lam = _sss_material(amount=0.0)
    sss = _sss_material(amount=1.0)
    f_lam = oracle.render(_plane(lam).arrays(), fr, NEE, 0, 32)["film"][..., 1].mean()
    f_sss = oracle.render(_plane(sss).arrays(), fr, NEE, 0, 32)["film"][..., 1].mean()
    want = float(_mixture_cdf(RMAX_FACTOR, 1.0))                             # the profile mass within r_max
    assert abs(f_sss / f_lam - want) < 0.01, (f_sss, f_lam)


def _lit_scene():
    b = scenes.MeshBuilder()
    floor = scenes.preset_material("PLASTIC")
    floor.albedo[:] = (0.7, 0.7, 0.7)
    floor.specular_amount = 0.0
    floor.roughness = 1.0
    skin = _sss_material(albedo=0.8, radius=0.15)
    skin.albedo[:] = (0.9, 0.6, 0.5)
    skin.subsurface_radius[:] = (0.3, 0.12, 0.06)
    skin.specular_amount = 0.5
    skin.roughness = 0.4
    b.quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), b.material(floor))
    P, N, I = scenes.sphere_template(16, 24)
    b.add(np.asarray(P) * 0.8 + [0, 0.8, 0], N, I, b.material(skin))
    bx, bn, bi = scenes.box_template()
    b.add(np.asarray(bx) * [0.5, 0.4, 0.5] + [1.4, 0.4, 0.3], bn, bi, b.material(skin))
    lm = scenes.blank_material()
    lm.emissive[:] = (12.0, 12.0, 12.0)
    b.quad((-0.6, 3.0, -0.6), (0, 0, 1.2), (1.2, 0, 0), b.material(lm))
    sun = _lib.Light(_lib.Light.DIRECTIONAL)
    sun.lookAt((3, 6, 2), (0, 0, 0))
    sun.setIntensity(1.5)
    cam = camera.Camera(eye=(0.5, 2.2, 4.5), target=(0.3, 0.6, 0), fovy=0.8)
    return b.arrays([sun]), cam


def test_bpt_and_nee_agree_on_a_subsurface_scene(oracle):
    scene, cam = _lit_scene()
    fr = camera.make_frame(cam, 40, 30, max_depth=6, firefly_clamp_scale=0.0)
    a = oracle.render(scene, fr, BPT, 0, 96)["film"][..., 1]
    b = oracle.render(scene, fr, NEE, 0, 96)["film"][..., 1]
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert abs(a.mean() / b.mean() - 1) < 0.03, (a.mean(), b.mean())


def test_bdpt_agrees_with_nee_on_a_subsurface_scene(oracle):
    """subsurface exits in both subpaths, with their area-density edges kept out of the MIS sums where
    no strategy can connect them: BDPT converges to the same image as NEE"""
    scene, cam = _lit_scene()
    fr = camera.make_frame(cam, 40, 30, max_depth=6, firefly_clamp_scale=0.0, light_path_vertices=6)
    a = oracle.render(scene, fr, NEE, 0, 384)["film"][..., 1]
    b = oracle.render(scene, fr, BDPT, 0, 384)["film"][..., 1]
    assert np.isfinite(b).all()
    assert abs(b.mean() / a.mean() - 1) < 0.015, (a.mean(), b.mean())
    # per 10 x 10 block with enough light, where one estimator's bias would show against the other's noise
    ba, bb = a.reshape(3, 10, 4, 10).mean((1, 3)), b.reshape(3, 10, 4, 10).mean((1, 3))
    lit = ba > 0.25 * ba.max()
    assert lit.sum() >= 6 and np.all(np.abs(bb[lit] / ba[lit] - 1) < 0.05), bb / ba


@pytest.mark.gpu
@pytest.mark.parametrize("integrator,wavefront", [(BPT, 1), (NEE, 1), (BDPT, 1), (BDPT, 0)])
def test_subsurface_render_matches_the_checker(ylib, oracle, integrator, wavefront):
    scene, cam = _lit_scene()
    frame = camera.make_frame(cam, 64, 48, max_depth=6, light_path_vertices=6)
    it = ylib.Integrator(integrator, 0)
    if integrator == BDPT:
        it.set_option("bdpt_wavefront", wavefront)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 3)
    g_last, g = it.last_samples(), it.film()
    it.close()
    ref = oracle.render(scene, frame, integrator, 0, 3)
    assert np.array_equal(g_last[..., 3].view(np.uint32), ref["last"][..., 3].view(np.uint32))
    # the two shadow rays a subsurface path can add in one bounce are applied atomically, so their
    # sum may round differently
    assert np.allclose(g[..., :3], ref["film"][..., :3], rtol=1e-4, atol=1e-6)
