"""Learned denoiser: the CPU checker pinned by finite differences, the BF16 matrix-core kernels against
a plain PyTorch FP32 convolution and the checker, the whole network's gradients against the checker,
and what training gains over the noisy input."""
import numpy as np
import pytest

from oracle import denoise_ref as R


def _fd_check(f, params, grads, idx_count=12, eps=1e-6, seed=0, rtol=1e-5):
    rng = np.random.default_rng(seed)
    for p, g in zip(params, grads):
        for _ in range(idx_count):
            i = tuple(int(rng.integers(s)) for s in p.shape)
            old = p[i]
            p[i] = old + eps
            fp = f()
            p[i] = old - eps
            fm = f()
            p[i] = old
            num = (fp - fm) / (2 * eps)
            assert abs(num - g[i]) <= rtol * max(1.0, abs(num)) + 1e-7, (i, num, g[i])


def test_bf16_rounding_matches_torch():
    import torch
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 3.0e38, 1 + 2 ** -8, 1 + 3 * 2 ** -8]])
    want = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(R.bf16(x), want)


def test_conv_gradients_by_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 5, 4, 3))
    w = rng.normal(size=(2, 3, 3, 3))
    b = rng.normal(size=2)
    Rw = rng.normal(size=(2, 5, 4, 2))
    f = lambda: float((R.conv3x3(x, w, b, relu=False) * Rw).sum())
    gx, gw, gb, _ = R.conv3x3_backward(x, w, R.conv3x3(x, w, b), Rw, relu=False)
    _fd_check(f, [x, w, b], [gx, gw, gb])


def test_pool_and_upsample_gradients_by_finite_differences():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 4, 6, 2))
    Rp = rng.normal(size=(1, 2, 3, 2))
    _fd_check(lambda: float((R.pool2(x) * Rp).sum()), [x], [R.pool2_grad(Rp)])
    small = rng.normal(size=(1, 2, 3, 2))
    skip = rng.normal(size=(1, 4, 6, 3))
    Ru = rng.normal(size=(1, 4, 6, 5))
    gin, gs = R.up2_concat_grad(Ru, 2)
    _fd_check(lambda: float((R.up2_concat(small, skip) * Ru).sum()), [small, skip], [gin, gs])


def test_unet_gradients_by_finite_differences():
    rng = np.random.default_rng(4)
    ch = (4, 6)
    shapes = [(4, 3, 3, 5), (4, 3, 3, 4), (6, 3, 3, 4), (6, 3, 3, 6), (4, 3, 3, 10), (4, 3, 3, 4), (5, 3, 3, 4)]
    params = [(rng.normal(0, 0.4, s), rng.normal(0, 0.1, s[0])) for s in shapes]
    x = rng.normal(size=(1, 4, 4, 5))
    Rw = rng.normal(size=(1, 4, 4, 5))

    def f():
        y, _ = R.unet_forward(x, params, ch, round_bf16=False)
        return float((y * Rw).sum())

    y, cache = R.unet_forward(x, params, ch, round_bf16=False)
    grads = R.unet_backward(Rw, params, ch, cache, round_bf16=False)
    _fd_check(f, [p[0] for p in params] + [p[1] for p in params], [g[0] for g in grads] + [g[1] for g in grads],
              idx_count=4, eps=1e-6, rtol=1e-4)


# ---------------------------------------------------------------- GPU
def _bits(a):
    import torch
    from yapt_amd import denoiser
    return denoiser.to_bf16_bits(torch.from_numpy(np.asarray(a, np.float32)).cuda())


def _f(t):
    from yapt_amd import denoiser
    return denoiser.bf16_to_float(t).cpu().numpy().astype(np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 12, 20, 16, 32), (1, 10, 66, 32, 64), (2, 16, 16, 48, 128), (1, 18, 70, 96, 32)])
def test_conv_with_pooled_output_equals_conv_then_pool(shape):
    # the pooled output written from the conv's tile is ydn_pool2 of its output, bit for bit, and the
    # output itself is ydn_conv3x3's
    import torch
    from yapt_amd import denoiser
    lib = denoiser.load()
    N, H, W, Cin, Cout = shape
    rng = np.random.default_rng(9)
    xg = _bits(rng.normal(size=(N, H, W, Cin)))
    wg = _bits(rng.normal(0, 0.2, size=(Cout, 3, 3, Cin)))
    bg = torch.from_numpy(rng.normal(0, 0.1, size=Cout).astype(np.float32)).cuda()
    P = denoiser._p
    for relu in (0, 1):
        out = denoiser.bf16_empty((N, H, W, Cout))
        denoiser.check(lib.ydn_conv3x3(P(xg), N, H, W, Cin, P(wg), P(bg), Cout, relu, P(out), None))
        pool = denoiser.bf16_empty((N, H // 2, W // 2, Cout))
        denoiser.check(lib.ydn_pool2(P(out), N, H, W, Cout, P(pool), None))
        out2 = denoiser.bf16_empty((N, H, W, Cout))
        pool2 = denoiser.bf16_empty((N, H // 2, W // 2, Cout))
        denoiser.check(lib.ydn_conv3x3_pool2(P(xg), N, H, W, Cin, P(wg), P(bg), Cout, relu, P(out2), P(pool2), None))
        torch.cuda.synchronize()
        assert torch.equal(out, out2)
        assert torch.equal(pool, pool2)
    bad = denoiser.bf16_empty((1, 4, 4, 32))
    assert lib.ydn_conv3x3_pool2(P(xg), 1, 5, 4, Cin, P(wg), None, 32, 0, P(bad), P(bad), None) != 0  # odd H


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 12, 20, 16, 32), (1, 9, 7, 32, 64), (2, 16, 16, 48, 128), (1, 5, 33, 96, 96),
                                   (1, 19, 70, 48, 32), (1, 17, 65, 224, 96)])
def test_conv_forward_matches_torch_fp32_and_checker(shape):
    import torch
    from yapt_amd import denoiser
    lib = denoiser.load()
    N, H, W, Cin, Cout = shape
    rng = np.random.default_rng(5)
    x = R.bf16(rng.normal(size=(N, H, W, Cin)))
    w = R.bf16(rng.normal(0, 0.2, size=(Cout, 3, 3, Cin)))
    b = rng.normal(0, 0.1, size=Cout).astype(np.float32)
    xg, wg = _bits(x), _bits(w)
    bg = torch.from_numpy(b).cuda()
    for relu in (0, 1):
        out = denoiser.bf16_empty((N, H, W, Cout))
        denoiser.check(lib.ydn_conv3x3(denoiser._p(xg), N, H, W, Cin, denoiser._p(wg), denoiser._p(bg), Cout, relu,
                                       denoiser._p(out), None))
        torch.cuda.synchronize()
        got = _f(out)
        ref_t = torch.nn.functional.conv2d(torch.from_numpy(x).permute(0, 3, 1, 2).float().cuda(),
                                           torch.from_numpy(w).permute(0, 3, 1, 2).float().cuda(), bg, padding=1)
        ref_t = ref_t.permute(0, 2, 3, 1).cpu().numpy().astype(np.float64)
        if relu:
            ref_t = np.maximum(ref_t, 0)
        ref_c = R.conv3x3(x, w, b.astype(np.float64), relu=bool(relu))
        scale = np.abs(ref_c).max()
        assert np.abs(ref_t - ref_c).max() < 1e-4 * scale                      # the checker is the torch op
        assert np.abs(got - ref_c).max() < 8e-3 * scale, np.abs(got - ref_c).max() / scale   # bf16 output rounding


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 12, 20, 32, 16, 32), (1, 10, 70, 64, 32, 64), (1, 6, 34, 128, 96, 96)])
def test_fused_upsample_concat_conv_is_the_two_step_conv(shape):
    # ydn_conv3x3_up2cat stages the upsampled coarse channels and the skip channels itself: the same values
    # reach the same matrix-core sums as ydn_up2_concat followed by ydn_conv3x3
    import torch
    from yapt_amd import denoiser
    lib = denoiser.load()
    N, H, W, C1, C2, Cout = shape
    rng = np.random.default_rng(9)
    coarse = _bits(R.bf16(rng.normal(size=(N, H // 2, W // 2, C1))))
    skip = _bits(R.bf16(rng.normal(size=(N, H, W, C2))))
    w = _bits(R.bf16(rng.normal(0, 0.2, size=(Cout, 3, 3, C1 + C2))))
    b = torch.from_numpy(rng.normal(0, 0.1, size=Cout).astype(np.float32)).cuda()
    cat = denoiser.bf16_empty((N, H, W, C1 + C2))
    two, one = denoiser.bf16_empty((N, H, W, Cout)), denoiser.bf16_empty((N, H, W, Cout))
    denoiser.check(lib.ydn_up2_concat(denoiser._p(coarse), C1, denoiser._p(skip), C2, N, H, W, denoiser._p(cat), None))
    denoiser.check(lib.ydn_conv3x3(denoiser._p(cat), N, H, W, C1 + C2, denoiser._p(w), denoiser._p(b), Cout, 1, denoiser._p(two), None))
    denoiser.check(lib.ydn_conv3x3_up2cat(denoiser._p(coarse), C1, denoiser._p(skip), C2, N, H, W, denoiser._p(w), denoiser._p(b),
                                          Cout, 1, denoiser._p(one), None))
    torch.cuda.synchronize()
    assert torch.equal(one, two)
    assert _f(one).max() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(2, 12, 20, 16, 32), (1, 10, 6, 32, 64), (1, 33, 17, 64, 32)])
def test_conv_gradients_match_checker(shape):
    import torch
    from yapt_amd import denoiser
    lib = denoiser.load()
    N, H, W, Cin, Cout = shape
    rng = np.random.default_rng(6)
    x = R.bf16(rng.normal(size=(N, H, W, Cin)))
    w = R.bf16(rng.normal(0, 0.2, size=(Cout, 3, 3, Cin)))
    g = R.bf16(rng.normal(size=(N, H, W, Cout)))
    gx_ref, gw_ref, gb_ref, _ = R.conv3x3_backward(x, w, None, g, relu=False)
    xg, wg, gg = _bits(x), _bits(w), _bits(g)
    dw = torch.zeros((Cout, 3, 3, Cin), dtype=torch.float32, device="cuda")
    db = torch.zeros(Cout, dtype=torch.float32, device="cuda")
    denoiser.check(lib.ydn_conv3x3_wgrad(denoiser._p(xg), denoiser._p(gg), N, H, W, Cin, Cout, denoiser._p(dw), denoiser._p(db), None))
    wt = denoiser.bf16_empty((Cin, 3, 3, Cout))
    denoiser.check(lib.ydn_flip_weights(denoiser._p(wg), Cout, Cin, denoiser._p(wt), None))
    torch.cuda.synchronize()
    if Cin % 32 == 0:
        gx = denoiser.bf16_empty((N, H, W, Cin))
        denoiser.check(lib.ydn_conv3x3(denoiser._p(gg), N, H, W, Cout, denoiser._p(wt), None, Cin, 0, denoiser._p(gx), None))
        torch.cuda.synchronize()
        assert np.abs(_f(gx) - gx_ref).max() < 8e-3 * np.abs(gx_ref).max()
    assert np.abs(dw.cpu().numpy() - gw_ref).max() < 1e-4 * np.abs(gw_ref).max()
    assert np.abs(db.cpu().numpy() - gb_ref).max() < 1e-4 * np.abs(gb_ref).max()


@pytest.mark.gpu
def test_unet_gradients_match_checker():
    import torch
    from yapt_amd import denoiser
    ch = (32, 32, 64)
    net = denoiser.UNet(channels=ch, seed=3)
    rng = np.random.default_rng(7)
    x = R.bf16(rng.normal(size=(2, 16, 12, denoiser.IN_CHANNELS)))
    noisy = np.abs(rng.normal(1, 0.5, size=(2, 16, 12, 4))).astype(np.float32)
    ref = np.abs(rng.normal(1, 0.2, size=(2, 16, 12, 4))).astype(np.float32)
    params = [(R.bf16(c.w.cpu().numpy()), c.b.cpu().numpy().astype(np.float64)) for c in net.layers()]
    y = net.forward(_bits(x), keep=True)
    loss, g, _ = denoiser.loss_and_grad(y, torch.from_numpy(noisy).cuda(), torch.from_numpy(ref).cuda())
    net.zero_grad()
    net.backward(g)
    torch.cuda.synchronize()
    y_ref, cache = R.unet_forward(x, params, ch)
    assert np.abs(_f(y) - y_ref).max() < 2e-2 * np.abs(y_ref).max()
    l_ref, g_ref = R.l1_log_loss(y_ref, noisy.astype(np.float64), ref.astype(np.float64))
    assert abs(float(loss.item()) - l_ref) < 1e-2 * l_ref
    grads = R.unet_backward(g_ref, params, ch, cache)
    for k, (c, (gw, gb)) in enumerate(zip(net.layers(), grads)):
        dw = c.dw.cpu().numpy()
        err = np.linalg.norm(dw - gw) / np.linalg.norm(gw)
        cos = float((dw * gw).sum() / (np.linalg.norm(dw) * np.linalg.norm(gw)))
        assert err < 0.05 and cos > 0.998, (k, err, cos)


@pytest.mark.gpu
def test_training_beats_the_noisy_input():
    import torch
    from yapt_amd import denoiser
    data = denoiser.make_dataset(8, res=128, spp_in=2, spp_ref=256, seed=5)
    train, held = data[:6], data[6:]
    net = denoiser.UNet(channels=(32, 32, 64, 64, 96), seed=1)
    hist = denoiser.train(net, train, iters=300, batch=4, crop=64, lr=2e-3, seed=2)
    assert hist[-1][1] < hist[0][1]
    gain = []
    for noisy, feats, ref in held:
        out = denoiser.denoise(net, torch.from_numpy(noisy).cuda(), torch.from_numpy(feats).cuda()).cpu().numpy()
        gain.append(R_rel(noisy, ref) / R_rel(out, ref))
    assert min(gain) > 1.5, gain


def R_rel(img, ref):
    from yapt_amd import denoiser
    return denoiser.rel_mse(img, ref)
