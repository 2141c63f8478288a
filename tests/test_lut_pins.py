"""Pins of the energy-compensation tables (MultipleScatterLUTUtility.h:200-438, MultipleScatterLUT.cpp:66-380)
and of the functions they are made from.

Each table is checked twice, with code written here in NumPy from the published methods and not from
the generator under test:
* against an independent quadrature of the integral it estimates (deterministic grids, exact
  Fresnel from Snell's law, Smith's height-correlated masking for GGX (Heitz 2014), the "Charlie"
  sheen distribution with the Estevez-Kulla 2017 Lambda fit). The reference's estimator is a
  256-point Hammersley rule (1024-point Halton for the translucent tables), whose first point sits
  at the interval's left end; its bias against the exact integral is a few 1e-3, which bounds the
  tolerance of these checks;
* against a NumPy replica of that estimator (Heitz 2018 visible-normal sampling, the same point
  sets), which must give the baked numbers to 3e-5.
A plausible mistake (masking without height correlation, alpha = r instead of r^2, a missing
Fresnel factor, swapped table axes, another point set) moves entries by 1e-2 or more.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SS, SSMS, TR, SH = 32, 16, 16, 16
MIN_COS = 0.001


@pytest.fixture(scope="module")
def tables(oracle):
    luts = oracle.ms_luts().astype(np.float64)
    out, o = {}, 0
    for name, shape in (("ss", (SS, SS)), ("ss_avg", (SS,)), ("ssms", (SSMS,) * 3), ("ssms_avg", (SSMS, SSMS)),
                        ("tr_denser", (TR,) * 3), ("tr_lighter", (TR,) * 3), ("tr_avg_denser", (TR, TR)),
                        ("tr_avg_lighter", (TR, TR)), ("sheen", (SH, SH))):
        n = int(np.prod(shape))
        out[name] = luts[o:o + n].reshape(shape)
        o += n
    assert o == len(luts)
    return out


@pytest.fixture(scope="module")
def consts():
    with open(os.path.join(HERE, "golden", "material_constants.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- physics, written from the published forms
def fresnel_exact(eta, c):
    """unpolarised reflectance of a dielectric interface, relative index eta = n_t / n_i, incident cosine c"""
    c = np.clip(np.asarray(c, np.float64), 0.0, 1.0)
    s2 = (1.0 - c * c) / (eta * eta)
    ct = np.sqrt(np.maximum(0.0, 1.0 - s2))
    with np.errstate(invalid="ignore", divide="ignore"):
        rs = (c - eta * ct) / (c + eta * ct)
        rp = (eta * c - ct) / (eta * c + ct)
    return np.where(s2 >= 1.0, 1.0, 0.5 * (rs * rs + rp * rp))


def lambda_ggx(cos, a):
    c2 = np.maximum(np.asarray(cos, np.float64) ** 2, 1e-300)
    return 0.5 * (np.sqrt(1.0 + a * a * (1.0 - c2) / c2) - 1.0)


def lambda_sheen(cos, r, k):
    """Estevez & Kulla 2017: L(x) = a / (1 + b x^c) + d x + e, coefficients lerped by (1 - r)^2 between the
    r = 1 and r = 0 fits; Lambda = exp(L(cos)) below cos 0.5, exp(2 L(0.5) - L(1 - cos)) above"""
    t = (1.0 - r) ** 2
    p = [k[n + "1"] * (1 - t) + k[n + "0"] * t for n in "abcde"]

    def L(x):
        return p[0] / (1.0 + p[1] * np.power(x, p[2])) + p[3] * x + p[4]
    cos = np.asarray(cos, np.float64)
    return np.exp(np.where(cos < 0.5, L(cos), 2.0 * L(0.5) - L(1.0 - cos)))


def d_sheen(cos_m, r):
    return (2.0 + 1.0 / r) * np.power(np.maximum(0.0, 1.0 - cos_m * cos_m), 0.5 / r) / (2.0 * math.pi)


def ndf_grid(a, n=800, m=256):
    """microfacet normals distributed as D(m) cos(theta_m) (GGX), on a midpoint grid in (u, phi) of
    tan^2 theta_m = a^2 u / (1 - u): an average over the grid integrates g / (D cos) over the hemisphere"""
    u = (np.arange(n) + 0.5) / n
    ct = 1.0 / np.sqrt(1.0 + a * a * u / (1.0 - u))
    phi = (np.arange(m) + 0.5) / m * 2.0 * math.pi
    ct, phi = np.meshgrid(ct, phi, indexing="ij")
    st = np.sqrt(1.0 - ct * ct)
    return np.stack([st * np.cos(phi), ct, st * np.sin(phi)])


def hemisphere_grid(n=400, m=128):
    """directions on a midpoint grid in (cos, phi): an average times 2 pi integrates over the hemisphere"""
    mu = (np.arange(n) + 0.5) / n
    phi = (np.arange(m) + 0.5) / m * 2.0 * math.pi
    mu, phi = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1.0 - mu * mu)
    return np.stack([s * np.cos(phi), mu, s * np.sin(phi)])


def wo_of(mu):
    return np.array([math.sqrt(1.0 - mu * mu), mu, 0.0])


def exact_reflection(mu, a, eta=None):
    """E(mu) = int f_r cos dw_i of GGX reflection with height-correlated masking, Fresnel of eta (None: 1)"""
    M = ndf_grid(a)
    wo = wo_of(mu)
    om = np.tensordot(wo, M, 1)
    wi = 2.0 * om[None] * M - wo[:, None, None]
    g2 = 1.0 / (1.0 + lambda_ggx(mu, a) + lambda_ggx(wi[1], a))
    F = 1.0 if eta is None else fresnel_exact(eta, om)
    return np.where((om > 0) & (wi[1] > 0), F * g2 * om / (mu * M[1]), 0.0).mean()


def exact_translucent(mu, a, eta):
    """reflection plus refraction through a rough dielectric interface (Walter et al. 2007), eta = n_t / n_i"""
    M = ndf_grid(a)
    wo = wo_of(mu)
    om = np.tensordot(wo, M, 1)
    F = fresnel_exact(eta, om)
    wr = 2.0 * om[None] * M - wo[:, None, None]
    k = 1.0 - (1.0 - om * om) / (eta * eta)
    wt = -wo[:, None, None] / eta + (om / eta - np.sqrt(np.maximum(k, 0.0)))[None] * M
    lo = lambda_ggx(mu, a)
    refl = np.where(wr[1] > 0, F / (1.0 + lo + lambda_ggx(wr[1], a)), 0.0)
    refr = np.where((k > 0) & (wt[1] < 0), (1.0 - F) / (1.0 + lo + lambda_ggx(wt[1], a)), 0.0)
    return np.where(om > 0, (refl + refr) * om / (mu * M[1]), 0.0).mean()


def exact_sheen(mu, r, k):
    a = max(max(0.01, r), 0.07)
    W = hemisphere_grid()
    wo = wo_of(mu)
    wh = W + wo[:, None, None]
    wh /= np.linalg.norm(wh, axis=0)
    g = 1.0 / (1.0 + lambda_sheen(W[1], a, k) + lambda_sheen(mu, a, k))
    f = g * d_sheen(wh[1], a) / (4.0 * mu * W[1])
    return (f * W[1]).mean() * 2.0 * math.pi


def sample_lut_cpu(table_row, u):
    """MultipleScatterLUTUtility.h sampleLut: linear between floor and ceil of u * dim, clamped to the last entry"""
    dim = len(table_row)
    v = np.minimum(np.clip(u, 0, 1) * dim, dim - 1)
    f = np.floor(v)
    c = np.ceil(v)
    t = v - f
    return (1 - t) * table_row[f.astype(int)] + t * table_row[c.astype(int)]


def sample_lut2_cpu(t2, ux, uy):
    dy = t2.shape[0]
    vy = min(np.clip(uy, 0, 1) * dy, dy - 1)
    y0, y1 = int(math.floor(vy)), int(math.ceil(vy))
    ty = vy - math.floor(vy)
    return (1 - ty) * sample_lut_cpu(t2[y0], ux) + ty * sample_lut_cpu(t2[y1], ux)


def exact_ssms(mu, r, eta, ss, ss_avg):
    """single scattering with Fresnel plus the Kulla-Conty lobe F(wo.m) (1 - E(mu_o)) (1 - E(mu_i)) / (pi (1 - E_avg)),
    E and E_avg read from the single-scattering tables with the CPU lookup"""
    r = max(0.01, r)
    single = exact_reflection(mu, r * r, eta)
    W = hemisphere_grid()
    wo = wo_of(mu)
    wh = W + wo[:, None, None]
    wh /= np.linalg.norm(wh, axis=0)
    F = fresnel_exact(eta, np.tensordot(wo, wh, 1))
    eo = 1.0 - sample_lut2_cpu(ss, mu, r)
    ei = 1.0 - np.array([sample_lut2_cpu(ss, x, r) for x in W[1, :, 0]])[:, None]
    eavg = sample_lut_cpu(ss_avg, np.array([r]))[0]
    ms = F * eo * ei / max(1e-4, (1.0 - eavg) * math.pi)
    return single + (ms * W[1]).mean() * 2.0 * math.pi


# ---------------------------------------------------------------- replicas of the reference's estimators
def van_der_corput(i):
    i = np.asarray(i, np.uint64)
    out = np.zeros_like(i, dtype=np.float64)
    f = 0.5
    for _ in range(32):
        out += (i & 1) * f
        i = i >> np.uint64(1)
        f *= 0.5
    return out


def hammersley(n):
    i = np.arange(n)
    return van_der_corput(i), i / n


def radical_inverse(base, i):
    i = np.asarray(i, np.int64).copy()
    out = np.zeros(len(i))
    f = 1.0 / base
    while np.any(i > 0):
        out += (i % base) * f
        i //= base
        f /= base
    return out


def vndf(wo, a, u1, u2):
    """Heitz 2018, "Sampling the GGX distribution of visible normals", isotropic, y-up local frame"""
    V = np.array([a * wo[0], a * wo[2], wo[1]])
    V /= np.linalg.norm(V)
    lensq = V[0] ** 2 + V[1] ** 2
    T1 = np.array([-V[1], V[0], 0.0]) / math.sqrt(lensq) if lensq > 0 else np.array([1.0, 0.0, 0.0])
    T2 = np.cross(V, T1)
    rr = np.sqrt(u1)
    phi = 2 * math.pi * u2
    t1, t2 = rr * np.cos(phi), rr * np.sin(phi)
    s = 0.5 * (1.0 + V[2])
    t2 = (1.0 - s) * np.sqrt(1.0 - t1 * t1) + s * t2
    N = t1[:, None] * T1 + t2[:, None] * T2 + np.sqrt(np.maximum(0.0, 1.0 - t1 * t1 - t2 * t2))[:, None] * V
    m = np.stack([a * N[:, 0], np.maximum(0.0, N[:, 2]), a * N[:, 1]], 1)
    return m / np.linalg.norm(m, axis=1, keepdims=True)


def estimate_reflection(mu, r, eta=None, kulla=None):
    """the reference's directional-albedo estimator: 256 Hammersley points, visible normals, weight G2 D (F) /
    (4 mu mu_i) |mu_i| / (D_v J)"""
    r = max(0.01, r)
    a = r * r
    wo = wo_of(mu)
    u1, u2 = hammersley(256)
    m = vndf(wo, a, u1, u2)
    om = m @ wo
    wi = 2.0 * om[:, None] * m - wo
    ok = (wi[:, 1] >= 0) & (om >= 0) if eta is not None else (wi[:, 1] > 0) & (om > 0)
    lo, li = lambda_ggx(mu, a), lambda_ggx(wi[:, 1], a)
    g2 = 1.0 / (1.0 + lo + li)
    g1 = 1.0 / (1.0 + lo)
    F = 1.0 if eta is None else fresnel_exact(eta, om)
    w = F * g2 / g1
    if kulla is not None:
        ss, ss_avg = kulla
        eo = 1.0 - sample_lut2_cpu(ss, mu, r)
        ei = 1.0 - np.array([sample_lut2_cpu(ss, x, r) for x in wi[:, 1]])
        eavg = sample_lut_cpu(ss_avg, np.array([r]))[0]
        ems = eo * ei / max(1e-4, (1.0 - eavg) * math.pi)
        # pdf of wi = D_v(m) / (4 (wi.m)) = G1 (wo.m) D / (mu 4 (wo.m)) ; |wi.y| / pdf of the compensation term
        D = a * a / (math.pi * (m[:, 1] ** 2 * (a * a - 1.0) + 1.0) ** 2)
        pdf = g1 * D / (4.0 * mu)
        w = w + F * ems * np.abs(wi[:, 1]) / np.where(pdf > 0, pdf, np.inf)
    return np.where(ok, w, 0.0).sum() / 256


def estimate_sheen(mu, r, k):
    r = max(0.01, r)
    a = max(r, 0.07)
    z, u2 = hammersley(256)
    rr = np.sqrt(1.0 - z * z)
    phi = 2 * math.pi * u2
    wi = np.stack([rr * np.cos(phi), z, rr * np.sin(phi)], 1)
    wo = wo_of(mu)
    wh = wi + wo
    wh /= np.linalg.norm(wh, axis=1, keepdims=True)
    g = 1.0 / (1.0 + lambda_sheen(wi[:, 1], a, k) + lambda_sheen(mu, a, k))
    f = g * d_sheen(wh[:, 1], a) / np.maximum(4.0 * mu * wi[:, 1], 1e-6)
    return np.where(wi[:, 1] > 0, f * np.abs(wi[:, 1]) * 2 * math.pi, 0.0).sum() / 256


def avg_albedo(row):
    """2 int E(mu) mu dmu for E linear between equally spaced nodes on [0, 1]"""
    h = 1.0 / (len(row) - 1)
    a = np.arange(len(row) - 1) * h
    b = a + h
    return 2.0 * np.sum(h * (row[:-1] * (2 * a + b) + row[1:] * (a + 2 * b)) / 6.0)


def coords(dim, step):
    return [(i, max(i / (dim - 1), MIN_COS)) for i in range(0, dim, step)]


# ---------------------------------------------------------------- tests
def test_single_scattering_table_is_the_ggx_albedo(tables):
    ss = tables["ss"]
    worst = 0.0
    for y, _ in coords(SS, 3):
        r = max(0.01, y / (SS - 1))
        for x, mu in coords(SS, 3):
            if mu < 0.09:
                continue  # grazing: the left-endpoint estimator and the grids both resolve it poorly
            worst = max(worst, abs(ss[y, x] - exact_reflection(mu, r * r)))
    assert worst < 6e-3, worst


@pytest.mark.parametrize("y", [0, 5, 13, 22, 31])
def test_single_scattering_table_replicates_the_estimator(tables, y):
    r = y / (SS - 1)
    for x, mu in coords(SS, 1):
        # at cos 0.001 with alpha <= 1e-3 the single-precision estimator of the reference loses ~1e-3 to
        # round-off in the visible-normal sample; everywhere else the replica agrees to ~1e-6
        tol = 2e-3 if (x == 0 and r < 0.05) else 3e-5
        assert abs(tables["ss"][y, x] - estimate_reflection(mu, r)) < tol, (y, x)


def test_average_albedo_tables(tables):
    for y in range(SS):
        assert abs(tables["ss_avg"][y] - avg_albedo(tables["ss"][y])) < 1e-5
    for z in range(SSMS):
        for y in range(SSMS):
            assert abs(tables["ssms_avg"][z, y] - avg_albedo(tables["ssms"][z, y])) < 1e-5
    for name in ("denser", "lighter"):
        t, av = tables["tr_" + name], tables["tr_avg_" + name]
        for z in range(TR):
            for y in range(TR):
                assert abs(av[z, y] - avg_albedo(t[z, y])) < 1e-5
    # and the average of the exact albedo, 2 int E mu dmu, by a fine rule
    for y in (4, 16, 28):
        r = max(0.01, y / (SS - 1))
        mus = (np.arange(24) + 0.5) / 24
        exact = 2.0 * np.mean([exact_reflection(m, r * r) * m for m in mus])
        assert abs(tables["ss_avg"][y] - exact) < 8e-3, (y, tables["ss_avg"][y], exact)


def _ssms_eta(z):
    return max(1.0 + 3.0 * z / (SSMS - 1), 1.01)


def test_fresnel_multiscatter_table_is_fresnel_albedo_plus_kulla_conty(tables):
    worst = 0.0
    for z in (0, 5, 15):
        eta = _ssms_eta(z)
        for y in (3, 9, 15):
            for x in (4, 9, 15):
                mu = x / (SSMS - 1)
                e = exact_ssms(mu, y / (SSMS - 1), eta, tables["ss"], tables["ss_avg"])
                worst = max(worst, abs(tables["ssms"][z, y, x] - e))
    assert worst < 5e-3, worst


def test_fresnel_multiscatter_table_replicates_the_estimator(tables):
    for z in (0, 7, 15):
        for y in (0, 6, 15):
            for x, mu in coords(SSMS, 3):
                e = estimate_reflection(mu, y / (SSMS - 1), _ssms_eta(z), kulla=(tables["ss"], tables["ss_avg"]))
                tol = 2e-3 if (x == 0 and y == 0) else 5e-5  # single-precision round-off at cos 0.001, alpha 1e-4
                assert abs(tables["ssms"][z, y, x] - e) < tol, (z, y, x, tables["ssms"][z, y, x], e)


def _tr_eta(name, z):
    if name == "denser":
        e = min(max(3.0 * z / (TR - 1) + 1.0, 1.0001), 4.0)
        return max(e, 1.01)
    e = min(max(0.69 * z / (TR - 1) + 0.3, 0.3), 0.99)
    return min(e, 0.99)


@pytest.mark.parametrize("name", ["denser", "lighter"])
def test_translucent_tables_are_the_rough_dielectric_albedo(tables, name):
    t = tables["tr_" + name]
    worst = 0.0
    for z in (0, 6, 15):
        eta = _tr_eta(name, z)
        for y in (2, 8, 15):
            a = max(0.01, y / (TR - 1)) ** 2
            for x in (3, 8, 15):
                worst = max(worst, abs(t[z, y, x] - exact_translucent(x / (TR - 1), a, eta)))
    assert worst < 1e-2, worst


@pytest.mark.parametrize("name", ["denser", "lighter"])
def test_translucent_tables_replicate_the_estimator(tables, name):
    # 1024 points of the Halton sequence in bases 2, 3, 5 with identity digit permutations (the reference
    # permutes digits with an unseeded rand(), which no table can reproduce): visible normal, then
    # reflection with probability F, else refraction; weight f |wi.y| / pdf
    n = 1024
    h = [np.minimum(radical_inverse(b, np.arange(n)), 1.0 - 2.0 ** -24) for b in (2, 3, 5)]
    t = tables["tr_" + name]
    for z in (0, 7, 15):
        eta = _tr_eta(name, z)
        for y in (0, 7, 15):
            a = max(0.01, y / (TR - 1)) ** 2
            for x, mu in coords(TR, 1):
                wo = wo_of(mu)
                m = vndf(wo, a, h[0], h[1])
                om = m @ wo
                F = fresnel_exact(eta, om)
                refl = h[2] < F
                wr = 2.0 * om[:, None] * m - wo
                k = 1.0 - (1.0 - om * om) / (eta * eta)
                wt = -wo / eta + (om / eta - np.sqrt(np.maximum(k, 0.0)))[:, None] * m
                wi = np.where(refl[:, None], wr, wt)
                ok = (om >= 0) & np.where(refl, wr[:, 1] > 0, (wt[:, 1] <= 0) & (k >= 0))
                g = (1.0 + lambda_ggx(mu, a)) / (1.0 + lambda_ggx(mu, a) + lambda_ggx(wi[:, 1], a))
                e = np.where(ok, g, 0.0).sum() / n
                tol = 2e-3 if (x == 0 and y == 0) else 1e-4
                assert abs(t[z, y, x] - e) < tol, (name, z, y, x, t[z, y, x], e)


def test_sheen_fit_coefficients_and_lambda(oracle, consts):
    k = consts["glambda_sheen"]
    # the fit Estevez & Kulla publish (Production Friendly Microfacet Sheen BRDF, 2017, table 1)
    assert (k["a0"], k["b0"], k["c0"], k["d0"], k["e0"]) == (25.3245, 3.32435, 0.16801, -1.27393, -4.85967)
    assert (k["a1"], k["b1"], k["c1"], k["d1"], k["e1"]) == (21.5473, 3.82987, 0.19823, -1.9776, -4.32054)
    rs = np.repeat([0.07, 0.3, 0.6, 1.0], 9)
    cs = np.tile(np.linspace(0.05, 0.95, 9), 4)
    got = oracle.evaluate("glambda_sheen", np.stack([rs, cs], 1).astype(np.float32))[:, 0]
    want = lambda_sheen(cs, rs, k)
    assert np.allclose(got, want, rtol=2e-5), np.abs(got / want - 1).max()


def test_sheen_distribution_is_normalised(oracle):
    # int D(m) cos(theta_m) dw_m = 1 for every roughness; the shader form (miscBrdf.hlsl:180-186) keeps
    # sin^2 theta_m at 1e-4 or more
    th = (np.arange(4000) + 0.5) / 4000 * (math.pi / 2)
    for r in (0.07, 0.2, 0.5, 1.0):
        d = oracle.evaluate("d_sheen", np.stack([np.cos(th), np.full_like(th, r)], 1).astype(np.float32))[:, 0]
        c = np.cos(th).astype(np.float32).astype(np.float64)  # the cosines the function received
        want = (2.0 + 1.0 / r) * np.power(np.maximum(1.0 - c * c, 1e-4), 0.5 / r) / (2.0 * math.pi)
        # near cos 1, 1 - cos^2 in single precision cancels: compare against the largest value there
        assert np.allclose(d, want, rtol=5e-5, atol=1e-6 * want.max())
        integral = (d * np.cos(th) * np.sin(th)).sum() * (math.pi / 2 / 4000) * 2 * math.pi
        assert abs(integral - 1.0) < 1e-3, (r, integral)


def test_sheen_table_is_the_sheen_albedo(tables, consts):
    k = consts["glambda_sheen"]
    worst = 0.0
    for y in (1, 5, 10, 15):
        for x in (2, 6, 11, 15):
            mu = x / (SH - 1)
            worst = max(worst, abs(tables["sheen"][y, x] - exact_sheen(mu, y / (SH - 1), k)))
            assert abs(tables["sheen"][y, x] - estimate_sheen(mu, y / (SH - 1), k)) < 3e-5, (y, x)
    assert worst < 2e-3, worst


def test_average_fresnel_fit(oracle, consts):
    f = consts["avg_fresnel"]
    etas = np.concatenate([np.linspace(0.35, 0.99, 12), np.linspace(1.01, 3.0, 12)])
    got = oracle.evaluate("avg_fresnel", etas[:, None].astype(np.float32))[:, 0]
    lo = f["below_1"]
    hi = f["above_1_denominator"]
    want = np.where(etas < 1, lo[0] + lo[1] * etas + lo[2] * etas ** 2 + lo[3] * etas ** 3, (etas - 1) / (hi[0] + hi[1] * etas))
    assert np.allclose(got, want, rtol=1e-5, atol=1e-7)
    # the fit approximates the hemispherical average 2 int F(mu) mu dmu of the exact Fresnel reflectance
    mus = (np.arange(20000) + 0.5) / 20000
    exact = np.array([2.0 * np.mean(fresnel_exact(e, mus) * mus) for e in etas])
    assert np.max(np.abs(got - exact)) < 1e-2, np.max(np.abs(got - exact))
