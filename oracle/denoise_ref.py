"""CPU checker for the denoiser kernels. TEST INFRASTRUCTURE ONLY (tests load it; the product never
does).

NumPy float64 restatements of the denoiser's operations - 3x3 convolution (zero padding, NHWC,
weights [Cout][3][3][Cin]) with bias and ReLU, its data/weight/bias gradients, 2x2 average pooling,
nearest 2x upsampling with concatenation, the log-space L1 loss - and of the whole U-Net forward and
backward built from them. `bf16` rounds to bfloat16 (round to nearest even) so the checker can round
at the same points as the kernels, which store every activation and gradient in bfloat16.
There is no reference implementation of this network (the reference denoiser is a bilateral
filter); the gradients here are pinned by central finite differences in the tests.
"""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def conv3x3(x, w, b=None, relu=False):
    N, H, W, C = x.shape
    xp = np.zeros((N, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    out = np.zeros((N, H, W, w.shape[0]))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + H, kx:kx + W], w[:, ky, kx, :], optimize=True)
    if b is not None:
        out += b
    return np.maximum(out, 0.0) if relu else out


def conv3x3_backward(x, w, y, g, relu=False):
    """gradients of the input, weights and bias given the layer output y and its gradient g"""
    N, H, W, C = x.shape
    if relu:
        g = g * (y > 0)
    xp = np.zeros((N, H + 2, W + 2, C))
    xp[:, 1:-1, 1:-1] = x
    gw = np.zeros_like(w, dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            gw[:, ky, kx, :] = np.einsum("nhwo,nhwc->oc", g, xp[:, ky:ky + H, kx:kx + W], optimize=True)
    gb = g.sum(axis=(0, 1, 2))
    gp = np.zeros((N, H + 2, W + 2, g.shape[3]))
    gp[:, 1:-1, 1:-1] = g
    gx = np.zeros((N, H, W, C))
    for ky in range(3):
        for kx in range(3):
            # out[h, w] uses x[h + ky - 1, w + kx - 1]  =>  gx[h', w'] += g[h' - ky + 1, w' - kx + 1] w[., ky, kx, .]
            gx += np.einsum("nhwo,oc->nhwc", gp[:, 2 - ky:2 - ky + H, 2 - kx:2 - kx + W], w[:, ky, kx, :], optimize=True)
    return gx, gw, gb, g


def pool2(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool2_grad(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up2_concat(x, skip):
    return np.concatenate([np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), skip], axis=3)


def up2_concat_grad(g, c1):
    N, H, W, _ = g.shape
    gin = g[..., :c1].reshape(N, H // 2, 2, W // 2, 2, c1).sum(axis=(2, 4))
    return gin, g[..., c1:]


def l1_log_loss(net, noisy, ref):
    """net (N, H, W, C) with the first 3 channels added to log(1 + noisy); mean |.| against log(1 + ref);
    returns the loss and its gradient with respect to net (per-image mean, summed over the batch)"""
    N, H, W, C = net.shape
    pred = np.log1p(np.maximum(noisy[..., :3], 0)) + net[..., :3]
    t = np.log1p(np.maximum(ref[..., :3], 0))
    d = pred - t
    loss = np.abs(d).mean(axis=(1, 2, 3)).mean()
    g = np.zeros_like(net)
    g[..., :3] = np.sign(d) / (3.0 * H * W)
    return loss, g


def unet_forward(x, params, channels, round_bf16=True):
    """params: list of (w, b) in the order encoder (2 per level), decoder (2 per level, level L-2 first),
    output. Returns the output and the cache of (input, output) per conv and the skips."""
    r = bf16 if round_bf16 else (lambda a: a)
    L = len(channels)
    k = 0
    cache = []
    skips = []
    h = x
    for l in range(L):
        if l > 0:
            h = r(pool2(h))
        for _ in range(2):
            w, b = params[k]
            y = r(conv3x3(h, w, b, relu=True))
            cache.append((h, y))
            h = y
            k += 1
        skips.append(h)
    for l in range(L - 2, -1, -1):
        h = up2_concat(h, skips[l])
        for _ in range(2):
            w, b = params[k]
            y = r(conv3x3(h, w, b, relu=True))
            cache.append((h, y))
            h = y
            k += 1
    w, b = params[k]
    y = r(conv3x3(h, w, b, relu=False))
    cache.append((h, y))
    return y, cache


def unet_backward(g, params, channels, cache, round_bf16=True):
    """weight and bias gradients (same order as params) of the U-Net given the output gradient g"""
    r = bf16 if round_bf16 else (lambda a: a)
    L = len(channels)
    grads = [None] * len(params)
    n_enc = 2 * L
    k = len(params) - 1

    def back(k, g, relu=True, need_x=True):
        x, y = cache[k]
        w, _ = params[k]
        gx, gw, gb, _ = conv3x3_backward(x, w, y, g, relu)
        grads[k] = (gw, gb)
        return r(gx) if need_x else None

    g = back(k, r(g), relu=False)
    k -= 1
    gskip = [None] * L
    for l in range(L - 1):
        g = back(k, g)
        g = back(k - 1, g)
        k -= 2
        gin, gs = up2_concat_grad(g, channels[l + 1])
        g, gskip[l] = r(gin), gs
    for l in range(L - 1, -1, -1):
        if l < L - 1:
            g = r(g + gskip[l])
        g = back(2 * l + 1, g)
        g = back(2 * l, g, need_x=l > 0)
        if l > 0:
            g = r(pool2_grad(g))
    assert k == n_enc - 1
    return grads
