"""Shared FP32 core functions against closed forms, published vectors and Monte Carlo identities."""
import math

import numpy as np
import pytest


def _u32(x):
    return np.asarray(x, dtype=np.uint64).astype(np.uint32)


# ---------------------------------------------------------------- sampler (random.hlsl)
def _fmix32(x):
    x ^= x >> 16
    x = (x * 0x85EBCA6B) & 0xFFFFFFFF
    x ^= x >> 13
    x = (x * 0xC2B2AE35) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def test_murmur_finaliser_published_vectors():
    # MurmurHash3_x86_32 of the empty input is fmix32(seed): ("", 1) = 0x514E28B7, ("", 0xffffffff) = 0x81F16F39
    assert _fmix32(1) == 0x514E28B7
    assert _fmix32(0xFFFFFFFF) == 0x81F16F39


def test_hash_pair(oracle):
    rng = np.random.default_rng(0)
    a = rng.integers(0, 1 << 32, size=2000, dtype=np.uint64)
    b = rng.integers(0, 600, size=2000, dtype=np.uint64)
    args = np.stack([_u32(a), _u32(b)], axis=1)
    got = oracle.evaluate("hash_pair", args)[:, 0].view(np.uint32)
    for i in range(len(a)):
        s = (int(a[i]) + int(b[i])) & 0xFFFFFFFF
        key = ((s * ((s + 1) & 0xFFFFFFFF)) & 0xFFFFFFFF) // (2 + int(b[i]))
        assert int(got[i]) == _fmix32(key)
    assert int(oracle.evaluate("hash_pair", np.array([[0, 0]], np.uint32))[0, 0].view(np.uint32)) == 0


def test_owen_scramble_preserves_stratification(oracle):
    # nested uniform scrambling keeps every elementary interval of a (0, m, 1)-sequence occupied once
    vdc = [int("{:032b}".format(i)[::-1], 2) for i in range(256)]
    rng = np.random.default_rng(1)
    for seed in rng.integers(0, 1 << 32, size=20, dtype=np.uint64):
        args = np.stack([_u32(vdc), _u32([seed] * 256)], axis=1)
        v = oracle.evaluate("owen_scramble", args)[:, 0].view(np.uint32).astype(np.uint64)
        for m in (1, 4, 8):
            n = 1 << m
            assert sorted((v[:n] >> (32 - m)).tolist()) == list(range(n))


def test_uint_to_float01_closed_form(oracle):
    vals = _u32([0, 1, 0x80000000, 0xFFFFFF7F, 0xFFFFFF80, 0xFFFFFFFF])
    got = oracle.evaluate("uint_to_float01", vals[:, None])[:, 0]
    assert got[0] == 0.0 and got[2] == 0.5
    assert got[-1] == np.float32(0.99999994) and got[-2] == np.float32(0.99999994)
    assert np.all(got < 1.0)


# ---------------------------------------------------------------- own transcendentals
@pytest.mark.parametrize("name,lo,hi,tol", [("sin", -60, 60, 2e-7), ("cos", -60, 60, 2e-7), ("exp", -80, 80, 3e-7),
                                            ("log", 1e-30, 1e4, 4e-7)])
def test_transcendentals_against_float64(oracle, name, lo, hi, tol):
    rng = np.random.default_rng(2)
    if name == "log":
        x = np.exp(rng.uniform(np.log(lo), np.log(hi), 50000)).astype(np.float32)
    else:
        x = rng.uniform(lo, hi, 50000).astype(np.float32)
    got = oracle.evaluate(name, x[:, None])[:, 0].astype(np.float64)
    ref = {"sin": np.sin, "cos": np.cos, "exp": np.exp, "log": np.log}[name](x.astype(np.float64))
    if name == "exp":
        err = np.abs(got - ref) / ref
    elif name == "log":
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)
    else:
        err = np.abs(got - ref)
    assert err.max() < tol


def test_transcendental_special_values(oracle):
    assert oracle.evaluate("exp", np.array([[0.0]]))[0, 0] == 1.0
    assert oracle.evaluate("log", np.array([[1.0]]))[0, 0] == 0.0
    assert oracle.evaluate("exp", np.array([[-200.0]]))[0, 0] == 0.0
    assert np.isinf(oracle.evaluate("exp", np.array([[100.0]]))[0, 0])
    assert oracle.evaluate("log", np.array([[0.0]]))[0, 0] == -np.inf
    assert oracle.evaluate("pow", np.array([[0.0, 2.0]]))[0, 0] == 0.0
    assert oracle.evaluate("pow", np.array([[0.3, 0.0]]))[0, 0] == 1.0
    p = oracle.evaluate("pow", np.array([[0.5, 3.0], [2.0, 10.0], [0.9, 0.5]]))[:, 0]
    assert np.allclose(p, [0.125, 1024.0, math.sqrt(0.9)], rtol=2e-6)


# ---------------------------------------------------------------- spectral sigmoid
def test_sigmoid_special_cases(oracle):
    got = oracle.evaluate("sigmoid", np.array([[np.inf], [-np.inf], [0.0]], np.float32))[:, 0]
    assert list(got) == [1.0, 0.0, 0.5]
    inv = oracle.evaluate("sigmoid_inv", np.array([[1.0], [0.0], [0.5]], np.float32))[:, 0]
    assert inv[0] == np.inf and inv[1] == -np.inf and inv[2] == 0.0
    x = np.linspace(0.01, 0.99, 99).astype(np.float32)
    back = oracle.evaluate("sigmoid", oracle.evaluate("sigmoid_inv", x[:, None])[:, :1])[:, 0]
    assert np.allclose(back, x, atol=2e-7)


# ---------------------------------------------------------------- Fresnel
def test_fresnel_dielectric_closed_forms(oracle):
    for eta in (1.33, 1.5, 2.4):
        f0 = oracle.evaluate("fresnel_dielectric", np.array([[eta, 1.0]], np.float32))[0, 0]
        assert abs(f0 - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
        assert abs(oracle.evaluate("fresnel_dielectric", np.array([[eta, 0.0]], np.float32))[0, 0] - 1.0) < 1e-5
    # total internal reflection from the dense side beyond the critical angle
    eta = 1 / 1.5
    crit = math.sqrt(1 - eta * eta)
    c = np.array([[eta, crit * 0.9], [eta, 0.1]], np.float32)
    assert np.all(oracle.evaluate("fresnel_dielectric", c)[:, 0] == 1.0)
    # against the textbook s/p form for an interior angle
    cos_i = 0.7
    sin_t = math.sqrt(1 - cos_i ** 2) / 1.5
    cos_t = math.sqrt(1 - sin_t ** 2)
    rs = (cos_i - 1.5 * cos_t) / (cos_i + 1.5 * cos_t)
    rp = (1.5 * cos_i - cos_t) / (1.5 * cos_i + cos_t)
    got = oracle.evaluate("fresnel_dielectric", np.array([[1.5, cos_i]], np.float32))[0, 0]
    assert abs(got - 0.5 * (rs * rs + rp * rp)) < 1e-6


def test_fresnel_conductor_normal_incidence(oracle):
    # at normal incidence the reference's form reduces to ((n-1)^2 + k^2) / ((n+1)^2 + k^2)
    for n, k in ((0.2, 3.0), (1.1, 2.5), (0.05, 4.2)):
        got = oracle.evaluate("fresnel_conductor", np.array([[n, k, 1.0]], np.float32))[0, 0]
        assert abs(got - ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)) < 2e-6


# ---------------------------------------------------------------- GGX
def _dirs(theta, phi):
    return np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1)


@pytest.mark.parametrize("ax,ay", [(0.3, 0.3), (0.6, 0.2), (0.9, 0.9)])
def test_ggx_normalisation(oracle, ax, ay):
    # int D(m) cos(theta_m) dm = 1 over the hemisphere (midpoint rule)
    nt, nph = 400, 256
    th = (np.arange(nt) + 0.5) * (np.pi / 2) / nt
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    m = _dirs(T, P).reshape(-1, 3)
    args = np.concatenate([m, np.tile([[ax, ay]], (len(m), 1))], 1)
    D = oracle.evaluate("d_ggx", args)[:, 0].reshape(T.shape)
    integral = (D * np.cos(T) * np.sin(T)).sum() * (np.pi / 2 / nt) * (2 * np.pi / nph)
    assert abs(integral - 1.0) < 5e-3


@pytest.mark.parametrize("theta_o", [0.0, 0.6, 1.2])
def test_vndf_sampling_matches_its_pdf(oracle, theta_o):
    # E[1 / pdf(m)] over sampled m equals the solid angle of the sampled support: the lune where m.y > 0
    # and wo.m > 0, of angle pi - theta_o, i.e. 2 (pi - theta_o)
    ax = ay = 0.5
    wo = np.array([math.sin(theta_o), math.cos(theta_o), 0.0], np.float32)
    rng = np.random.default_rng(4)
    n = 200000
    u = rng.random((n, 2)).astype(np.float32)
    args = np.concatenate([np.tile(wo, (n, 1)), np.tile([[ax, ay]], (n, 1)), u], 1)
    m = oracle.evaluate("sample_wm_ggx", args)[:, :3]
    pdf = oracle.evaluate("pdf_wm_ggx", np.concatenate([np.tile(wo, (n, 1)), m, np.tile([[ax, ay]], (n, 1))], 1))[:, 0]
    ok = pdf > 0
    est = np.mean(np.where(ok, 1.0 / np.where(ok, pdf, 1.0), 0.0))
    # the estimator has heavy tails near the support boundary; compare with a loose bound
    assert ok.mean() > 0.999
    assert abs(est - 2 * (math.pi - theta_o)) / (2 * (math.pi - theta_o)) < 0.05


def test_cosine_hemisphere_samples(oracle):
    rng = np.random.default_rng(5)
    u = rng.random((100000, 2)).astype(np.float32)
    w = oracle.evaluate("sample_cosine_hemisphere", u)[:, :3]
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-6)
    assert abs(w[:, 1].mean() - 2.0 / 3.0) < 3e-3    # E[cos] under cos/pi
    assert abs(w[:, 0].mean()) < 5e-3 and abs(w[:, 2].mean()) < 5e-3


def test_emission_profile_conserves_power(oracle):
    # (n+2)/2 |cos|^n integrated against cos over the hemisphere equals pi for every n
    th = (np.arange(2000) + 0.5) * (np.pi / 2) / 2000
    for n in (0.0, 2.0, 8.0, 32.0):
        prof = oracle.evaluate("emission_profile", np.stack([np.cos(th), np.full_like(th, n)], 1).astype(np.float32))[:, 0]
        integral = (prof * np.cos(th) * np.sin(th)).sum() * (np.pi / 2 / 2000) * 2 * np.pi
        assert abs(integral - math.pi) / math.pi < 2e-3


# ---------------------------------------------------------------- energy-compensation tables
def test_directional_albedo_tables(oracle):
    luts = oracle.ms_luts()
    ss = luts[:32 * 32].reshape(32, 32)
    assert np.all(ss > 0) and np.all(ss <= 1.0 + 1e-3)
    assert np.all(ss[0, 1:] > 0.97)                   # near-smooth GGX reflects everything (no Fresnel)
    assert np.all(np.diff(ss[:, -1]) <= 1e-3)         # rougher -> less single scattering at normal incidence
    avg = luts[1024:1056]
    # integrateAverageAlbedo: Eavg = 2 int E(mu) mu dmu, bounded by the extremes of E
    assert np.all(avg <= ss.max(axis=1) + 1e-4) and np.all(avg >= ss.min(axis=1) * 0.99)
    ssms = luts[1056:1056 + 4096]
    assert np.all(ssms >= 0) and np.all(ssms < 1.05)
    sheen = luts[-256:]
    assert np.all(sheen >= 0)


# ---------------------------------------------------------------- RGB -> spectrum fit
@pytest.mark.parametrize("rgb", [(0.63, 0.065, 0.05), (0.14, 0.45, 0.091), (0.2, 0.3, 0.9), (0.8, 0.8, 0.1)])
def test_rgb_fit_reproduces_colour(oracle, rgb):
    grid = np.round(np.asarray(rgb) * 63) / 63  # a grid node: the fit there is exact up to the solver's tolerance
    c = oracle.rgb_coeffs(grid.astype(np.float32))
    back = oracle.rgb_reproduce(c)
    assert np.allclose(back, grid, atol=1e-5), (back, grid)
    # between nodes the trilinear table still gives a close colour
    c2 = oracle.rgb_coeffs(np.asarray(rgb, np.float32))
    assert np.allclose(oracle.rgb_reproduce(c2), rgb, atol=2.5e-2)


def test_rgb_fit_round_trip_at_grid_nodes(oracle):
    # the 64^3 coefficient table reproduces its node colours under D65 (SpectralUtility.cpp:465-609 solves
    # to 1e-6 in CIELAB^2); the fit's round-trip error stays far below the 1e-3 bar on random nodes
    rng = np.random.default_rng(5)
    worst = 0.0
    for n in rng.integers(0, 64, size=(400, 3)):
        g = (n / 63).astype(np.float32)
        worst = max(worst, float(np.abs(oracle.rgb_reproduce(oracle.rgb_coeffs(g)) - g).max()))
    assert worst < 1e-5, worst
