"""Train the denoiser on the tracer's own renders and judge it at 1920x1080.

    python -m yapt_amd.denoise_bench [--scenes 2000] [--iters 6000] [--out yapt_amd/data/denoiser_weights.npz]

1. Dataset: 2,000 seeded random-shapes scenes at 128x128, NEE, 4 spp inputs (film + first-hit
   features) and 1024 spp targets.
2. Training: Adam, random 128x128 crops, batch 8, cosine learning-rate decay.
3. Judgement on a held-out scene (the 1M-triangle random-shapes scene, seed 2, 1920x1080): relMSE
   against a 4096-spp plain render, for plain 64 spp and for 4 spp + denoiser, with the wall time of
   each (render, and render + denoise, on the GPU).
4. The denoiser's conv kernels: achieved TFLOP/s of one 1080p forward pass (HIP events).
5. The reference's non-learned denoiser (feature prefilter + 5 a-trous passes) on the same 4-spp
   film, as the comparison point: relMSE in linear sRGB for every variant.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


XYZ_TO_SRGB = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])


def xyz_to_linear_srgb(xyz):
    """XYZ film -> linear sRGB with the gamut clamp the FinalColor path applies"""
    rgb = np.asarray(xyz, np.float64)[..., :3] @ XYZ_TO_SRGB.T
    lum = rgb @ np.array([0.2126, 0.7152, 0.0722])
    m = rgb.min(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(lum - m != 0, lum / np.where(lum - m != 0, lum - m, 1), 0.0)
    return np.where((m < 0)[..., None], lum[..., None] + t[..., None] * (rgb - lum[..., None]), rgb)


def conv_flops(net, H, W):
    """multiply-add FLOPs of one forward pass at H x W (2 per multiply-add)"""
    total = 0
    L = len(net.ch)
    for l in range(L):
        P = (H >> l) * (W >> l)
        for c in net.enc[l]:
            total += 2 * P * 9 * c.cin * c.cout
    for l in range(L - 1):
        P = (H >> l) * (W >> l)
        for c in net.dec[l]:
            total += 2 * P * 9 * c.cin * c.cout
    total += 2 * H * W * 9 * net.out.cin * net.out.cout
    return total


def load_weights(path=None):
    """the trained U-Net (yapt_amd/data/denoiser_weights.npz by default)"""
    from . import denoiser
    net = denoiser.UNet(seed=0)
    st = np.load(path or os.path.join(HERE, "data", "denoiser_weights.npz"))
    net.load_state({"layers": [{"w": st["w%d" % k], "b": st["b%d" % k]} for k in range(len(net.layers()))]})
    return net


def judge(net, W=1920, H=1080, ref_spp=4096, log=lambda m: None):
    """Held-out judgement on the 1M-triangle random-shapes scene (seed 2) at W x H: relMSE against a
    ref_spp plain render for plain 64 spp and for 4 spp + the denoiser, the wall time of each (render,
    and render + denoise, on the GPU), the reference's bilateral denoiser on the same 4-spp film, and
    the conv kernels' achieved TFLOP/s over one forward pass."""
    import torch
    from . import _lib, camera, denoiser, scenes, tuning
    scene, cam = scenes.random_shapes(n_shapes=2000, seed=2)
    frame = camera.make_frame(cam, W, H, max_depth=12)
    it = _lib.Integrator(_lib.YPT_INTEGRATOR_NEE, 0)
    for k, v in tuning.best_config("nee-shapes2000-%dx%d-d12" % (W, H)).items():
        it.set_option(k, v)
    it.set_option("first_hit_features", 1)
    it.set_scene(scene)
    it.set_frame(frame)
    it.render(0, 4)   # warm-up
    it.synchronize()
    log("reference %d spp" % ref_spp)
    it.clear_film()
    for s0 in range(0, ref_spp, 256):
        it.render(10000 + s0, 256)
        it.synchronize()
    ref = it.film()

    it.clear_film()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it.render(0, 64)
    it.synchronize()
    t_plain = time.perf_counter() - t0
    plain = it.film()

    it.clear_film()
    x_noisy = x_feat = None
    denoiser.denoise(net, torch.from_numpy(plain).cuda(), torch.from_numpy(it.first_hit_features()).cuda())  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    it.render(0, 4)
    it.synchronize()
    t_render4 = time.perf_counter() - t0
    noisy = it.film()
    feats = it.first_hit_features()
    t1 = time.perf_counter()
    x_noisy = torch.from_numpy(noisy).cuda()
    x_feat = torch.from_numpy(feats).cuda()
    out = denoiser.denoise(net, x_noisy, x_feat)
    torch.cuda.synchronize()
    t_denoise = time.perf_counter() - t1
    den = out.cpu().numpy()
    # the reference's own (non-learned) denoiser on the same 4-spp film: prefilter + 5 a-trous passes
    bil_params = _lib.PostParams.defaults(denoise=1, tonemap=0)
    it.final_color(bil_params, linear=True)   # warm-up
    t2 = time.perf_counter()
    _, bil = it.final_color(bil_params, linear=True)
    t_bil = time.perf_counter() - t2

    # conv kernels: achieved TFLOP/s over one forward pass at the evaluation size (device time)
    x = denoiser.pack_input(x_noisy[None], x_feat[None], net.multiple)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    net.forward(x)
    e0.record()
    reps = 5
    for _ in range(reps):
        net.forward(x)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    flops = conv_flops(net, x.shape[1], x.shape[2])

    r_plain = denoiser.rel_mse(plain, ref)
    r_noisy = denoiser.rel_mse(noisy, ref)
    r_den = denoiser.rel_mse(den, ref)
    ref_s = xyz_to_linear_srgb(ref)
    srgb = {"plain_64spp": denoiser.rel_mse(xyz_to_linear_srgb(plain), ref_s),
            "plain_4spp": denoiser.rel_mse(xyz_to_linear_srgb(noisy), ref_s),
            "learned_4spp": denoiser.rel_mse(xyz_to_linear_srgb(den), ref_s),
            "bilateral_4spp": denoiser.rel_mse(bil[..., :3], ref_s)}
    t_den_total = t_render4 + t_denoise
    res = {
        "metric": "relMSE vs %d-spp plain render, 4 spp + denoiser vs plain 64 spp (%dx%d)" % (ref_spp, W, H),
        "relmse": {"plain_64spp": r_plain, "plain_4spp": r_noisy, "denoised_4spp": r_den},
        "wall_s": {"plain_64spp": t_plain, "render_4spp": t_render4, "denoise": t_denoise, "denoised_4spp_total": t_den_total,
                   "bilateral_denoise": t_bil},
        "relmse_linear_srgb": srgb,
        "speedup_vs_plain_64spp": t_plain / t_den_total,
        "denoised_matches_64spp": r_den <= r_plain,
        "gain_over_4spp": r_noisy / r_den,
        "conv_roofline": {"bound": "mfma", "achieved": flops / (fwd_ms * 1e-3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                          "frac": flops / (fwd_ms * 1e-3) / 1e12 / 2500.0, "forward_ms": fwd_ms, "flops": flops},
        "channels": list(net.ch),
    }
    it.close()
    return res


def forward_roofline(net, W=1920, H=1080, reps=20, warmup=5, seed=0):
    """the conv kernels' achieved TFLOP/s over one U-Net forward pass at W x H (padded to the network's
    multiple), timed with HIP events on the library's stream over `reps` passes after `warmup` untimed
    ones (the first passes run while the GPU's clocks ramp up), on a seeded input"""
    import torch
    from . import denoiser
    m = net.multiple
    Hp, Wp = (H + m - 1) // m * m, (W + m - 1) // m * m
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = denoiser.to_bf16_bits(torch.randn((1, Hp, Wp, denoiser.IN_CHANNELS), device="cuda", generator=gen) * 0.5)
    for _ in range(warmup):
        net.forward(x)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        net.forward(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = conv_flops(net, Hp, Wp)
    return {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
            "frac": flops / (ms * 1e-3) / 1e12 / 2500.0, "forward_ms": ms, "flops": flops, "shape": [Hp, Wp]}


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--forward-only", action="store_true",
                   help="time the U-Net forward pass at the evaluation size and print its conv roofline")
    p.add_argument("--scenes", type=int, default=2000)
    p.add_argument("--res", type=int, default=128)
    p.add_argument("--spp-ref", type=int, default=1024)
    p.add_argument("--iters", type=int, default=6000)
    p.add_argument("--batch", type=int, default=8)
    p.add_argument("--crop", type=int, default=128)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--eval-width", type=int, default=1920)
    p.add_argument("--eval-height", type=int, default=1080)
    p.add_argument("--eval-ref-spp", type=int, default=4096)
    p.add_argument("--out", default=os.path.join(HERE, "data", "denoiser_weights.npz"))
    p.add_argument("--load", default=None, help="skip training and load these weights")
    a = p.parse_args()
    import torch
    from . import denoiser

    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.forward_only:
        net = load_weights(a.load)
        print(json.dumps(forward_roofline(net, a.eval_width, a.eval_height)))
        return
    net = denoiser.UNet(seed=0)
    t_data = t_train = 0.0
    hist = []
    if a.load:
        net = load_weights(a.load)
    else:
        t0 = time.time()
        data = denoiser.make_dataset(a.scenes, res=a.res, spp_in=4, spp_ref=a.spp_ref, seed=100, log=log)
        t_data = time.time() - t0
        t0 = time.time()
        hist = denoiser.train(net, data, iters=a.iters, batch=a.batch, crop=a.crop, lr=a.lr, seed=1, log=log)
        torch.cuda.synchronize()
        t_train = time.time() - t0
        st = net.state()
        np.savez(a.out, **{"w%d" % k: l["w"] for k, l in enumerate(st["layers"])},
                 **{"b%d" % k: l["b"] for k, l in enumerate(st["layers"])}, channels=np.array(net.ch))
        del data

    res = judge(net, a.eval_width, a.eval_height, a.eval_ref_spp, log)
    res["training"] = {"scenes": a.scenes, "res": a.res, "spp_ref": a.spp_ref, "iters": a.iters, "batch": a.batch,
                       "crop": a.crop, "data_s": t_data, "train_s": t_train,
                       "loss_first_last": [hist[0][1], hist[-1][1]] if hist else None}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
