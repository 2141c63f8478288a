"""CPU checker for the steps after the film. TEST INFRASTRUCTURE ONLY (tests load it; the product never
does).

NumPy restatement, vectorized over the image, of:
  * the feature-moment accumulation of the film merge (mergeNewSamples.comp:93-113,
    denoiseFeatures.hlsl:29-44),
  * featuresFromMoments and the half-precision feature packing (denoiseFeatures.hlsl:63-146),
  * the feature prefilter (prefilterFeatures.comp:50-162),
  * one edge-avoiding a-trous pass (denoise.comp:142-270) with reconstructWorldPosition (:104-115),
  * XYZ -> linear sRGB with gamutClamp (colorSpaces.hlsl, denoise.comp:272-282), the ACES fitted
    operator (tonemapUtils.hlsl:1-36) and the sRGB transfer function with 8-bit quantization.
Float64 throughout, except the half-precision packing, which rounds as the kernels do.
"""
import numpy as np

HALF_MAX = 65000.0
G3 = np.array([3 / 8, 1 / 4, 1 / 16])
G3x3 = np.array([1 / 4, 1 / 8, 1 / 16])      # by |dx| + |dy|
XYZ_TO_SRGB = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])
ACES_IN = np.array([[0.59719, 0.35458, 0.04823], [0.07600, 0.90834, 0.01566], [0.02840, 0.13383, 0.83777]])
ACES_OUT = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])


def feature_moments(rec):
    """rec: (..., 16) per-sample feature records -> (4, ..., 4) moments (zeros for misses)"""
    rec = np.asarray(rec, np.float64)
    hit = (rec[..., 11] > 0) & np.all(np.isfinite(rec[..., :13]), axis=-1)
    a, n = rec[..., 3:6], rec[..., 0:3]
    mat = np.stack([rec[..., 6], rec[..., 8], rec[..., 9], rec[..., 12]], -1)
    ior = rec[..., 7] * 0.5
    d = rec[..., 10]
    m = np.stack([
        np.concatenate([a, (a * a).sum(-1, keepdims=True)], -1),
        np.concatenate([n, np.ones_like(d)[..., None]], -1),
        mat,
        np.stack([ior, (mat * mat).sum(-1) + ior * ior, d, d * d], -1),
    ])
    return np.where(hit[None, ..., None], m, 0.0)


def accumulate(records):
    """running means over passes as the merge keeps them: first pass copied, then lerp by 1/(n+1)"""
    acc = None
    for k, rec in enumerate(records):
        m = feature_moments(rec)
        acc = m if k == 0 else acc + (m - acc) * (1.0 / (k + 1))
    return acc


def features_from_moments(m):
    l0, l1, l2, l3 = m
    f = {"albedo": l0[..., :3], "normal": l1[..., :3], "mat": l2.copy(), "ior": l3[..., 0]}
    f["varAlbedo"] = np.maximum(l0[..., 3] - (f["albedo"] ** 2).sum(-1), 0)
    f["varMaterial"] = np.maximum(l3[..., 1] - (l2 ** 2).sum(-1) - f["ior"] ** 2, 0)
    cov = l1[..., 3]
    with np.errstate(divide="ignore", invalid="ignore"):
        depth = np.where(cov > 0, l3[..., 2] / np.where(cov > 0, cov, 1), 0.0)
        vd = np.where(cov > 0, np.maximum(l3[..., 3] / np.where(cov > 0, cov, 1) - depth ** 2, 0), 0.0)
    f["depth"], f["varDepth"] = depth, vd
    return f


def _h(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float16).astype(np.float64)


def pack_unpack(f):
    """write + read of the packed features: half precision everywhere but the depth (float32)"""
    depth = np.asarray(f["depth"], np.float32).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(f["depth"] > 0, np.minimum(f["varDepth"] / np.where(f["depth"] > 0, f["depth"] ** 2, 1), HALF_MAX), 0.0)
    g = {"albedo": _h(f["albedo"]), "normal": _h(f["normal"]), "mat": _h(f["mat"]), "ior": _h(f["ior"]),
         "varAlbedo": _h(np.minimum(f["varAlbedo"], HALF_MAX)), "varMaterial": _h(np.minimum(f["varMaterial"], HALF_MAX)),
         "depth": depth}
    g["varDepth"] = _h(rel) * depth * depth
    return g


def _vw(d2, sigma, vs, scale):
    return np.exp(-d2 / (sigma * sigma + scale * vs + 1e-8))


def _shift(a, dx, dy):
    """a[y + dy, x + dx] for every (y, x), with a validity mask"""
    H, W = a.shape[:2]
    out = np.zeros_like(a)
    valid = np.zeros((H, W), bool)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    if (ys.stop - ys.start) > 0 and (xs.stop - xs.start) > 0:
        out[yd, xd] = a[ys, xs]
        valid[yd, xd] = True
    return out, valid


def _take(f, dx, dy):
    g, valid = {}, None
    for k, v in f.items():
        g[k], valid = _shift(v, dx, dy)
    return g, valid


def _normal_var(f):
    return np.clip(1 - np.linalg.norm(f["normal"], axis=-1), 0, 1)


def feature_weight(a, b, ns, ms, scale):
    wn = _vw(((a["normal"] - b["normal"]) ** 2).sum(-1), ns, _normal_var(a) + _normal_var(b), scale)
    wa = _vw(((a["albedo"] - b["albedo"]) ** 2).sum(-1), ms, a["varAlbedo"] + b["varAlbedo"], scale)
    wm = _vw(((a["mat"] - b["mat"]) ** 2).sum(-1) + (a["ior"] - b["ior"]) ** 2, ms, a["varMaterial"] + b["varMaterial"], scale)
    return wn * wa * wm


def _rel_depth_w(c, s, dss, scale):
    d = c["depth"] - s["depth"]
    w = _vw(d * d, c["depth"] * dss, c["varDepth"] + s["varDepth"], scale)
    return np.where((c["depth"] <= 0) | (s["depth"] <= 0), 1.0, w)


def prefilter(m, ns=0.06, ms=0.05, dss=0.02, vscale=1.0, enabled=True):
    c = features_from_moments(m)
    if not enabled:
        return pack_unpack(c)
    H, W = c["depth"].shape
    sums = {"albedo": np.zeros((H, W, 3)), "normal": np.zeros((H, W, 3)), "mat": np.zeros((H, W, 4)), "ior": np.zeros((H, W))}
    sw = np.zeros((H, W))
    sd = np.zeros((H, W))
    sdw = np.zeros((H, W))
    sva = np.zeros((H, W))
    svm = np.zeros((H, W))
    svd = np.zeros((H, W))
    svw = np.zeros((H, W))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            s, valid = _take(c, dx, dy)
            kw = G3[abs(dx)] * G3[abs(dy)]
            w = kw * feature_weight(c, s, ns, ms, vscale) * _rel_depth_w(c, s, dss, vscale) * valid
            for k in sums:
                sums[k] = sums[k] + (s[k] * (w[..., None] if s[k].ndim == 3 else w))
            sw += w
            hasd = s["depth"] > 0
            sd += np.where(hasd, s["depth"] * w, 0)
            sdw += np.where(hasd, w, 0)
            if abs(dx) <= 1 and abs(dy) <= 1:
                wv = G3x3[abs(dx) + abs(dy)] * feature_weight(c, s, ns, ms, 0.0) * _rel_depth_w(c, s, dss, 0.0) * valid
                sva += s["varAlbedo"] * wv
                svm += s["varMaterial"] * wv
                svd += np.where(hasd, s["varDepth"], c["varDepth"]) * wv
                svw += wv
    inv = 1 / np.maximum(sw, 1e-6)
    r = dict(c)
    r["albedo"] = sums["albedo"] * inv[..., None]
    r["normal"] = sums["normal"] * inv[..., None]
    r["mat"] = sums["mat"] * inv[..., None]
    r["ior"] = sums["ior"] * inv
    with np.errstate(divide="ignore", invalid="ignore"):
        r["depth"] = np.where((c["depth"] > 0) & (sdw > 0), sd / np.where(sdw > 0, sdw, 1), c["depth"])
    ivw = 1 / np.maximum(svw, 1e-6)
    r["varAlbedo"], r["varMaterial"], r["varDepth"] = sva * ivw, svm * ivw, svd * ivw
    return pack_unpack(r)


def world_positions(depth, uv_to_view, view_to_world):
    H, W = depth.shape
    yy, xx = np.mgrid[0:H, 0:W]
    uv = np.stack([(xx + 0.5) / W, (yy + 0.5) / H, np.zeros_like(depth), np.ones_like(depth)], -1)
    near = uv @ np.asarray(uv_to_view, np.float64).reshape(4, 4).T
    near = near[..., :3] / near[..., 3:4]
    pvs = near * (depth / np.maximum(np.abs(near[..., 2]), 1e-8))[..., None]
    pvs4 = np.concatenate([pvs, np.ones_like(depth)[..., None]], -1)
    return (pvs4 @ np.asarray(view_to_world, np.float64).reshape(4, 4).T)[..., :3]


def atrous(src, var, f, p, uv_to_view, view_to_world, ns=0.06, ms=0.05, dss=0.02, cs=3.0, vscale=1.0):
    """one a-trous pass: returns (colour, variance)"""
    src = np.asarray(src, np.float64)
    var = np.asarray(var, np.float64)
    H, W = var.shape
    fv = np.zeros((H, W))
    fw = np.zeros((H, W))
    for dy in range(-1, 2):
        for dx in range(-1, 2):
            v, valid = _shift(var, dx, dy)
            w = np.array([[1 / 16, 1 / 8, 1 / 16], [1 / 8, 1 / 4, 1 / 8], [1 / 16, 1 / 8, 1 / 16]])[dx + 1, dy + 1]
            fv += v * w * valid
            fw += w * valid
    std = np.sqrt(np.maximum(fv / np.maximum(fw, 1e-6), 1e-6))
    pos = world_positions(f["depth"], uv_to_view, view_to_world)
    nl = np.linalg.norm(f["normal"], axis=-1)
    step = 1 << p
    sc4 = np.zeros((H, W, 4))
    sv = np.zeros((H, W))
    sw = np.zeros((H, W))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            s, valid = _take(f, dx * step, dy * step)
            scol, _ = _shift(src, dx * step, dy * step)
            svar, _ = _shift(var, dx * step, dy * step)
            spos, _ = _shift(pos, dx * step, dy * step)
            kw = G3[abs(dx)] * G3[abs(dy)]
            with np.errstate(divide="ignore", invalid="ignore"):
                dist = np.where(nl > 1e-3, ((f["normal"] / np.maximum(nl, 1e-30)[..., None]) * (spos - pos)).sum(-1),
                                f["depth"] - s["depth"])
            wd = np.where(s["depth"] <= 0, 1.0, _vw(dist * dist, f["depth"] * dss, f["varDepth"] + s["varDepth"], vscale))
            wm = feature_weight(f, s, ns, ms, vscale)
            wc = np.exp(-np.abs(scol[..., 1] - src[..., 1]) / (cs * std + 1e-6))
            w = kw * wd * wm * wc * valid
            sc4 += scol * w[..., None]
            sv += w * w * svar
            sw += w
    inv = 1 / np.maximum(sw, 1e-6)
    keep = f["depth"] <= 0
    col = np.where(keep[..., None], src, sc4 * inv[..., None])
    return col, np.where(keep, 0.0, sv * inv * inv)


def xyz_to_srgb(c):
    c = np.asarray(c, np.float64)
    rgb = c[..., :3] @ XYZ_TO_SRGB.T
    lum = rgb @ np.array([0.2126, 0.7152, 0.0722])
    m = rgb.min(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(lum - m != 0, lum / np.where(lum - m != 0, lum - m, 1), 0.0)
    clamped = lum[..., None] + t[..., None] * (rgb - lum[..., None])
    rgb = np.where((m < 0)[..., None], clamped, rgb)
    return np.concatenate([rgb, c[..., 3:4]], -1)


def aces_fitted(c):
    v = np.asarray(c, np.float64) @ ACES_IN.T
    a = v * (v + 0.0245786) - 0.000090537
    b = v * (0.983729 * v + 0.4329510) + 0.238081
    return np.clip((a / b) @ ACES_OUT.T, 0, 1)


def srgb_encode(c):
    c = np.clip(np.nan_to_num(np.asarray(c, np.float64), nan=0.0), 0, 1)
    e = np.where(c <= 0.0031308, 12.92 * c, 1.055 * np.power(c, 1 / 2.4) - 0.055)
    return np.floor(e * 255 + 0.5).astype(np.uint8)


def final_color(linear_rgba, exposure=1.0, tonemap=True):
    c = np.asarray(linear_rgba, np.float64) * exposure
    rgb = aces_fitted(c[..., :3]) if tonemap else c[..., :3]
    alpha = np.floor(np.clip(c[..., 3], 0, 1) * 255 + 0.5).astype(np.uint8)
    return np.concatenate([srgb_encode(rgb), alpha[..., None]], -1)
