"""The steps after the film: the CPU checker's closed forms (sRGB transfer, XYZ -> sRGB of the D65 white,
ACES on grey, gamut clamp), properties of its feature prefilter and a-trous denoiser, and on the GPU the
FinalColor readback, the feature moments and the denoiser against the checker."""
import numpy as np
import pytest

from oracle import post_ref as R

IDENT = np.eye(4).reshape(-1)


def test_srgb_transfer_known_values():
    # IEC 61966-2-1: linear 0.5 encodes to 0.7354 (188/255), the linear segment ends at 0.0031308
    got = R.srgb_encode(np.array([0.0, 1.0, 0.5, 0.0031308, 0.18, 2.0, -1.0, np.nan]))
    assert list(got) == [0, 255, 188, 10, 118, 255, 0, 0]


def test_xyz_to_srgb_maps_d65_white_to_white():
    white = R.xyz_to_srgb(np.array([0.95047, 1.0, 1.08883, 1.0]))
    assert np.allclose(white[:3], 1.0, atol=2e-3) and white[3] == 1.0


def test_gamut_clamp_keeps_luminance_and_removes_negatives():
    rng = np.random.default_rng(0)
    xyz = np.concatenate([rng.uniform(0, 1, (500, 3)), np.ones((500, 1))], -1)
    rgb = R.xyz_to_srgb(xyz)
    raw = xyz[:, :3] @ R.XYZ_TO_SRGB.T
    lum = raw @ np.array([0.2126, 0.7152, 0.0722])
    assert np.allclose(rgb[:, :3] @ np.array([0.2126, 0.7152, 0.0722]), lum)
    neg = raw.min(1) < 0
    assert neg.any() and np.all(rgb[neg, :3].min(1) > -1e-12)
    assert np.allclose(rgb[~neg, :3], raw[~neg])


def test_aces_keeps_grey_grey_and_is_monotonic_and_bounded():
    x = np.linspace(0, 20, 400)
    g = R.aces_fitted(np.stack([x, x, x], -1))
    # both matrices have unit row sums up to the rounding of their published entries (1e-5)
    assert np.allclose(g[:, 0], g[:, 1], atol=1e-12) and np.allclose(g[:, 1], g[:, 2], atol=2e-5)
    assert np.all(np.diff(g[:, 1]) >= 0) and g.min() >= 0 and g.max() <= 1
    # RRT+ODT fit on grey: (v(v + 0.0245786) - 0.000090537) / (v(0.983729 v + 0.432951) + 0.238081)
    assert abs(g[20, 1] - (x[20] * (x[20] + 0.0245786) - 0.000090537) / (x[20] * (0.983729 * x[20] + 0.4329510) + 0.238081)) < 1e-12


def _records(H, W, albedo, depth=5.0, normal=(0, 0, 1.0)):
    rec = np.zeros((H, W, 16))
    rec[..., 0:3] = normal
    rec[..., 3:6] = albedo
    rec[..., 6] = 0.5   # roughness
    rec[..., 7] = 1.5   # ior
    rec[..., 10] = depth
    rec[..., 11] = 1.0
    return rec


def _flat_film(H, W, value, rng=None, sigma=0.0):
    c = np.full((H, W, 4), value, np.float64)
    c[..., 3] = 1
    if rng is not None:
        c[..., :3] *= np.exp(rng.normal(0, sigma, (H, W, 1)))
    return c


def _denoise(film, var, moments, passes=5):
    f = R.prefilter(moments)
    col, v = film, var
    for p in range(passes):
        col, v = R.atrous(col, v, f, p, IDENT, IDENT)
    return col


def test_feature_moments_give_means_and_variances():
    H, W = 4, 5
    a = _records(H, W, (0.2, 0.4, 0.6), depth=2.0)
    b = _records(H, W, (0.6, 0.4, 0.2), depth=4.0)
    f = R.features_from_moments(R.accumulate([a, b]))
    assert np.allclose(f["albedo"], (0.4, 0.4, 0.4)) and np.allclose(f["depth"], 3.0)
    assert np.allclose(f["varAlbedo"], 2 * 0.2 ** 2) and np.allclose(f["varDepth"], 1.0)
    miss = np.zeros((H, W, 16))
    g = R.features_from_moments(R.accumulate([a, miss]))   # half the samples hit: depth is the mean over hits
    assert np.allclose(g["depth"], 2.0) and np.allclose(g["normal"][..., 2], 0.5)


def test_denoiser_keeps_a_constant_image():
    H, W = 24, 32
    film = _flat_film(H, W, 0.3)
    m = R.accumulate([_records(H, W, (0.5, 0.5, 0.5))])
    out = _denoise(film, np.full((H, W), 0.01), m)
    assert np.allclose(out, film, rtol=1e-12)


def test_denoiser_removes_noise_on_a_flat_region():
    rng = np.random.default_rng(1)
    H, W = 48, 48
    film = _flat_film(H, W, 0.3, rng, 0.3)
    m = R.accumulate([_records(H, W, (0.5, 0.5, 0.5))])
    out = _denoise(film, np.full((H, W), 0.09), m)
    err_in = ((film[..., 1] - 0.3 * np.exp(0.045)) ** 2).mean()
    err_out = ((out[..., 1] - 0.3 * np.exp(0.045)) ** 2).mean()
    assert err_out < 0.1 * err_in, (err_in, err_out)


def test_denoiser_does_not_blur_across_an_albedo_edge():
    rng = np.random.default_rng(2)
    H, W = 32, 32
    film = _flat_film(H, W, 0.2, rng, 0.2)
    film[:, W // 2:, :3] *= 3.0
    rec = _records(H, W, (0.2, 0.2, 0.2))
    rec[:, W // 2:, 3:6] = 0.8
    out = _denoise(film, np.full((H, W), 0.04), R.accumulate([rec]))
    left, right = out[:, : W // 2 - 1, 1].mean(), out[:, W // 2 + 1:, 1].mean()
    assert abs(left / film[:, : W // 2 - 1, 1].mean() - 1) < 0.02
    assert abs(right / film[:, W // 2 + 1:, 1].mean() - 1) < 0.02


def test_pixels_without_geometry_pass_through():
    H, W = 16, 16
    rng = np.random.default_rng(3)
    film = _flat_film(H, W, 0.5, rng, 0.5)
    out = _denoise(film, np.full((H, W), 0.25), R.accumulate([np.zeros((H, W, 16))]))
    assert np.array_equal(out, film)


def test_difference_stats_of_identical_noisy_and_shifted_captures():
    from yapt_amd.compare import difference_stats
    rng = np.random.default_rng(4)
    base = rng.uniform(40, 200, (1, 1, 3))
    a = np.clip(base + rng.normal(0, 20, (176, 240, 3)), 0, 255).astype(np.uint8)
    b = np.clip(base + rng.normal(0, 20, (176, 240, 3)), 0, 255).astype(np.uint8)
    assert difference_stats(a, a)["rel_down"] == 0
    s = difference_stats(a, b)   # two noisy captures of one image: the block average removes the noise
    assert s["rel_full"] > 10 * s["rel_down"]
    c = np.clip(a.astype(int) + 10, 0, 255).astype(np.uint8)
    assert difference_stats(a, c)["mean_gap"] > 0


# ---------------------------------------------------------------- GPU
def _render(ylib, scene, cam, spp, w=96, h=64, env=0.0, features=True, integrator=1, begin=0, depth=6):
    from yapt_amd import camera
    it = ylib.Integrator(integrator, 0)
    it.set_option("first_hit_features", 1 if features else 0)
    it.set_scene(scene)
    frame = camera.make_frame(cam, w, h, max_depth=depth, env_intensity_scale=env)
    it.set_frame(frame)
    it.render(begin, spp)
    return it, frame


@pytest.mark.gpu
def test_final_color_matches_the_checker(ylib, cornell):
    it, _ = _render(ylib, cornell[0], cornell[1], 8)
    film = it.film()
    for tonemap, exposure in ((1, 1.0), (0, 0.7)):
        p = ylib.PostParams.defaults(tonemap=tonemap, exposure=exposure)
        rgba, lin = it.final_color(p, linear=True)
        want_lin = R.xyz_to_srgb(film)
        assert np.allclose(lin, want_lin, rtol=1e-5, atol=1e-6)
        want = R.final_color(lin, exposure, bool(tonemap))
        d = np.abs(rgba.astype(int) - want.astype(int))
        assert d.max() <= 1 and (d > 0).mean() < 1e-3, ((d > 0).mean(), d.max())
    it.close()


@pytest.mark.gpu
def test_feature_moments_accumulate_like_the_film(ylib, shapes_small):
    scene, cam = shapes_small
    from yapt_amd import camera
    it = ylib.Integrator(ylib.YPT_INTEGRATOR_NEE, 0)
    it.set_option("first_hit_features", 1)
    it.set_scene(scene)
    it.set_frame(camera.make_frame(cam, 64, 48, max_depth=4))
    recs = []
    for k in range(5):
        it.render(k, 1)
        f = it.first_hit_features()
        rec = np.zeros(f.shape[:2] + (16,), np.float64)
        rec[..., :12] = f   # the test scenes have no anisotropy: record slot 12 is 0
        recs.append(rec)
    got = it.feature_moments()
    it.close()
    want = R.accumulate(recs)
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("integrator", [1, 2])
def test_denoiser_matches_the_checker(ylib, integrator):
    from yapt_amd import scenes
    scene, cam, env = scenes.punctual_lights("mixed", sphere_res=(12, 12))
    it, frame = _render(ylib, scene, cam, 4, w=80, h=56, env=env, integrator=integrator)
    film, var, mom = it.film(), it.film_variance(), it.feature_moments()
    _, lin = it.final_color(ylib.PostParams.defaults(denoise=1), linear=True)
    it.close()
    f = R.prefilter(mom.astype(np.float64))
    col, v = film.astype(np.float64), var.astype(np.float64)
    for p in range(5):
        col, v = R.atrous(col, v, f, p, np.asarray(frame.uv_to_view), np.asarray(frame.view_to_world))
    want = R.xyz_to_srgb(col)
    d = np.abs(lin[..., :3] - want[..., :3])
    bad = d > 2e-3 * np.maximum(np.abs(want[..., :3]), 1e-2)
    assert bad.mean() < 2e-3, (bad.mean(), d.max())


@pytest.mark.gpu
def test_bilateral_denoiser_lowers_the_error_of_a_noisy_render(ylib):
    from yapt_amd import denoiser, scenes
    scene, cam, env = scenes.punctual_lights("mixed", sphere_res=(16, 16))
    ref, _ = _render(ylib, scene, cam, 512, w=128, h=80, env=env, begin=1000)
    want = R.xyz_to_srgb(ref.film())[..., :3]
    ref.close()
    noisy, _ = _render(ylib, scene, cam, 4, w=128, h=80, env=env)
    _, plain = noisy.final_color(linear=True)
    _, den = noisy.final_color(ylib.PostParams.defaults(denoise=1), linear=True)
    noisy.close()
    e_plain = denoiser.rel_mse(plain[..., :3], want)
    e_den = denoiser.rel_mse(den[..., :3], want)
    assert e_den < 0.5 * e_plain, (e_plain, e_den)
