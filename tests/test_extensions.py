"""Extensions the benchmark text names and the reference does not have, both off by default: Sellmeier
dispersion (checked against the published BK7 indices) and blackbody emitters (checked against
Wien's displacement law, the Rayleigh-Jeans limit and the CIE 1931 Planckian locus through a render).
On the GPU a scene with BK7 glass and two blackbody emitters matches the checker for all three
integrators."""
import numpy as np
import pytest

from yapt_amd import _lib, camera, scenes

NEE, BPT, BDPT = _lib.YPT_INTEGRATOR_NEE, _lib.YPT_INTEGRATOR_BPT, _lib.YPT_INTEGRATOR_BDPT
(B, C) = scenes.BK7_SELLMEIER


def test_sellmeier_reproduces_the_published_bk7_indices(oracle):
    # Schott N-BK7: n_F (486.13 nm) 1.52238, n_d (587.56 nm) 1.51680, n_C (656.27 nm) 1.51432
    rows = [[lam, *B, *C] for lam in (486.13, 587.56, 656.27)]
    n = oracle.evaluate("sellmeier_ior", rows)[:, 0]
    assert np.allclose(n, [1.52238, 1.51680, 1.51432], atol=2e-5), n


def test_planck_spectrum_follows_wien_and_rayleigh_jeans(oracle):
    lam = np.arange(380.0, 700.0, 0.25)
    for T in (4500.0, 5000.0, 6000.0):
        f = oracle.evaluate("planck_relative", np.stack([lam, np.full_like(lam, T)], 1))[:, 0]
        peak = lam[np.argmax(f)]
        assert abs(peak - 2.897771955e6 / T) < 0.5, (T, peak)       # Wien's displacement law
    # far above the visible peak the spectrum tends to lambda^-4
    f = oracle.evaluate("planck_relative", [[380.0, 1e7], [700.0, 1e7]])[:, 0]
    assert abs(f[0] / f[1] / (700.0 / 380.0) ** 4 - 1) < 2e-3
    # normalised to 1 at 560 nm
    assert abs(oracle.evaluate("planck_relative", [[560.0, 3000.0]])[0, 0] - 1) < 1e-6


def _emitter_view(kelvin):
    b = scenes.MeshBuilder()
    m = scenes.blank_material()
    m.emissive[:] = (1.0, 1.0, 1.0)
    m.blackbody_kelvin = kelvin
    b.quad((-1, -1, 0), (2, 0, 0), (0, 2, 0), b.material(m))
    cam = camera.Camera(eye=(0, 0, 1), target=(0, 0, 0), fovy=0.5)
    return b.arrays(), cam


@pytest.mark.parametrize("kelvin,xy", [(2700.0, (0.4599, 0.4106)), (6500.0, (0.3135, 0.3237))])
def test_blackbody_emitter_lies_on_the_planckian_locus(oracle, kelvin, xy):
    scene, cam = _emitter_view(kelvin)
    fr = camera.make_frame(cam, 16, 16, max_depth=1)
    f = oracle.render(scene, fr, NEE, 0, 256)["film"][..., :3].astype(np.float64).sum((0, 1))
    x, y = f[0] / f.sum(), f[1] / f.sum()
    # CIE 1931 2-degree Planckian locus; the spectral range here ends at 700 nm, not 780
    assert abs(x - xy[0]) < 0.01 and abs(y - xy[1]) < 0.01, (x, y)


def _extension_scene():
    b = scenes.MeshBuilder()
    floor = scenes.preset_material("PLASTIC")
    floor.albedo[:] = (0.7, 0.7, 0.7)
    floor.specular_amount = 0.0
    floor.roughness = 1.0
    glass = scenes.preset_material("GLASS")
    glass.material_mask |= _lib.YPT_MATERIAL_DISPERSION
    glass.sellmeier_b[:] = B
    glass.sellmeier_c[:] = C
    warm, cool = scenes.blank_material(), scenes.blank_material()
    warm.emissive[:] = (8.0, 8.0, 8.0)
    warm.blackbody_kelvin = 2700.0
    cool.emissive[:] = (6.0, 6.0, 6.0)
    cool.blackbody_kelvin = 9000.0
    b.quad((-4, 0, 4), (8, 0, 0), (0, 0, -8), b.material(floor))
    P, N, I = scenes.sphere_template(16, 24)
    b.add(np.asarray(P) * 0.8 + [0, 0.8, 0], N, I, b.material(glass))
    b.quad((-2.0, 3.0, -0.5), (1.0, 0, 0), (0, 0, 1.0), b.material(warm))     # facing down
    b.quad((1.0, 3.0, -0.5), (1.0, 0, 0), (0, 0, 1.0), b.material(cool))
    return b.arrays(), camera.Camera(eye=(0.3, 2.0, 4.5), target=(0, 0.7, 0), fovy=0.8)


def test_extension_scene_renders_on_the_checker(oracle):
    scene, cam = _extension_scene()
    fr = camera.make_frame(cam, 32, 24, max_depth=6, light_path_vertices=6)
    for integ in (BPT, NEE, BDPT):
        f = oracle.render(scene, fr, integ, 0, 2)["film"]
        assert np.isfinite(f).all() and f[..., 1].max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [BPT, NEE, BDPT])
def test_extension_render_matches_the_checker(ylib, oracle, integrator):
    scene, cam = _extension_scene()
    frame = camera.make_frame(cam, 64, 48, max_depth=6, light_path_vertices=6)
    it = ylib.Integrator(integrator, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 3)
    g_last, g = it.last_samples(), it.film()
    it.close()
    ref = oracle.render(scene, frame, integrator, 0, 3)
    assert np.array_equal(g_last[..., 3].view(np.uint32), ref["last"][..., 3].view(np.uint32))
    assert np.allclose(g[..., :3], ref["film"][..., :3], rtol=1e-4, atol=1e-6)
