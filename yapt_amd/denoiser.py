"""Learned denoiser: a 5-level convolutional U-Net over the noisy film and first-hit features,
trained on renders of the path tracer itself, with every forward and backward operation running in
the project's BF16 matrix-core kernels (libyapt_denoise.so, include/yapt_denoise.h).

Torch tensors are used only as device memory (bf16 activations are stored as int16 tensors) and for
the stream; no torch operator computes any part of the network.

Network: input = [log(1 + colour), albedo, normal, 7 zero channels] (16 channels). Encoder level l
(l = 0..4, resolution / 2^l): two 3x3 conv + ReLU with widths 32, 64, 64, 96, 128, 2x2 average
pooling between levels (written by the level's last conv from its output tile). Decoder level l (3..0): nearest 2x upsample, concatenation with the encoder
output of level l, two 3x3 conv + ReLU. Output: 3x3 conv to 32 channels (the first 3 used), added
to log(1 + colour); loss L1 in that log space against a high-spp render.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libyapt_denoise.so")
CHANNELS = (32, 64, 64, 96, 128)
IN_CHANNELS = 16
OUT_CHANNELS = 32

_LIB = None


class DenoiseError(RuntimeError):
    pass


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise DenoiseError("libyapt_denoise.so is not built; run yapt_amd.build.build() (no CPU fallback exists)")
    from . import _lib
    _lib.share_hip_runtime()   # one HIP runtime for torch and the library
    lib = ctypes.CDLL(LIB_PATH)
    P, I, L, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
    lib.ydn_conv3x3.argtypes = [P, I, I, I, I, P, P, I, I, P, P]
    lib.ydn_conv3x3_wgrad.argtypes = [P, P, I, I, I, I, I, P, P, P]
    lib.ydn_conv3x3_up2cat.argtypes = [P, I, P, I, I, I, I, P, P, I, I, P, P]
    lib.ydn_conv3x3_pool2.argtypes = [P, I, I, I, I, P, P, I, I, P, P, P]
    lib.ydn_flip_weights.argtypes = [P, I, I, P, P]
    lib.ydn_relu_grad.argtypes = [P, P, L, P]
    lib.ydn_pool2.argtypes = [P, I, I, I, I, P, P]
    lib.ydn_pool2_grad.argtypes = [P, I, I, I, I, P, P]
    lib.ydn_up2_concat.argtypes = [P, I, P, I, I, I, I, P, P]
    lib.ydn_up2_concat_grad.argtypes = [P, I, I, I, I, I, P, P, I, P]
    lib.ydn_add.argtypes = [P, P, L, P]
    lib.ydn_pack_input.argtypes = [P, P, L, I, P, P]
    lib.ydn_l1_log_loss.argtypes = [P, I, P, P, L, P, P, P, P]
    lib.ydn_adam.argtypes = [P, P, P, P, L, F, F, F, F, I, P, P]
    lib.ydn_to_bf16.argtypes = [P, P, L, P]
    lib.ydn_last_error.restype = ctypes.c_char_p
    _LIB = lib
    return lib


def check(rc):
    if rc != 0:
        raise DenoiseError("yapt_denoise error %d: %s" % (rc, load().ydn_last_error().decode()))


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def bf16_empty(shape):
    import torch
    return torch.empty(shape, dtype=torch.int16, device="cuda")


def to_bf16_bits(t):
    """float tensor -> int16 tensor holding bfloat16 bits (round to nearest even, in the kernel)"""
    t = t.contiguous().float()
    out = bf16_empty(t.shape)
    check(load().ydn_to_bf16(_p(t), _p(out), t.numel(), _stream()))
    return out


def bf16_to_float(t):
    import torch
    return (t.to(torch.int32) << 16).view(torch.float32) if t.dtype == torch.int16 else t


# ---------------------------------------------------------------- layers
class Conv:
    """3x3 conv (+ReLU): FP32 master weights [Cout][3][3][Cin], BF16 copy, FP32 gradients, Adam state."""

    def __init__(self, cin, cout, relu, gen):
        import torch
        self.cin, self.cout, self.relu = cin, cout, relu
        std = math.sqrt(2.0 / (9 * cin))
        w = torch.from_numpy(gen.normal(0.0, std, size=(cout, 3, 3, cin)).astype(np.float32)).cuda()
        self.w = w
        self.b = torch.zeros(cout, dtype=torch.float32, device="cuda")
        self.w16 = to_bf16_bits(w)
        self.wt16 = bf16_empty((cin, 3, 3, cout))
        self.dw = torch.zeros_like(self.w)
        self.db = torch.zeros_like(self.b)
        self.mw, self.vw = torch.zeros_like(self.w), torch.zeros_like(self.w)
        self.mb, self.vb = torch.zeros_like(self.b), torch.zeros_like(self.b)
        self.x = self.y = None

    def forward(self, x, keep):
        N, H, W, C = x.shape
        assert C == self.cin, (C, self.cin)
        y = bf16_empty((N, H, W, self.cout))
        check(load().ydn_conv3x3(_p(x), N, H, W, C, _p(self.w16), _p(self.b), self.cout, int(self.relu), _p(y), _stream()))
        if keep:
            self.x, self.y = x, y
        return y

    def forward_pool(self, x, keep):
        """forward() and the 2x2 average pooling of its output, both written by the conv kernel"""
        N, H, W, C = x.shape
        assert C == self.cin, (C, self.cin)
        y = bf16_empty((N, H, W, self.cout))
        p = bf16_empty((N, H // 2, W // 2, self.cout))
        check(load().ydn_conv3x3_pool2(_p(x), N, H, W, C, _p(self.w16), _p(self.b), self.cout, int(self.relu), _p(y), _p(p),
                                       _stream()))
        if keep:
            self.x, self.y = x, y
        return y, p

    def forward_up2cat(self, coarse, skip):
        """the conv of [nearest 2x upsample of coarse, skip] without writing the concatenation (inference)"""
        N, H, W, c2 = skip.shape
        c1 = coarse.shape[3]
        assert c1 + c2 == self.cin, (c1, c2, self.cin)
        y = bf16_empty((N, H, W, self.cout))
        check(load().ydn_conv3x3_up2cat(_p(coarse), c1, _p(skip), c2, N, H, W, _p(self.w16), _p(self.b), self.cout,
                                        int(self.relu), _p(y), _stream()))
        return y

    def backward(self, g, need_input_grad=True):
        """g: gradient of this layer's output (bf16, overwritten); returns the input gradient."""
        lib = load()
        N, H, W, _ = g.shape
        s = _stream()
        if self.relu:
            check(lib.ydn_relu_grad(_p(self.y), _p(g), g.numel(), s))
        check(lib.ydn_conv3x3_wgrad(_p(self.x), _p(g), N, H, W, self.cin, self.cout, _p(self.dw), _p(self.db), s))
        if not need_input_grad:
            return None
        check(lib.ydn_flip_weights(_p(self.w16), self.cout, self.cin, _p(self.wt16), s))
        gx = bf16_empty((N, H, W, self.cin))
        check(lib.ydn_conv3x3(_p(g), N, H, W, self.cout, _p(self.wt16), None, self.cin, 0, _p(gx), s))
        return gx

    def zero_grad(self):
        self.dw.zero_()
        self.db.zero_()

    def step(self, lr, t, b1=0.9, b2=0.999, eps=1e-8):
        lib = load()
        s = _stream()
        check(lib.ydn_adam(_p(self.w), _p(self.dw), _p(self.mw), _p(self.vw), self.w.numel(), lr, b1, b2, eps, t, _p(self.w16), s))
        check(lib.ydn_adam(_p(self.b), _p(self.db), _p(self.mb), _p(self.vb), self.b.numel(), lr, b1, b2, eps, t, None, s))

    def params(self):
        return {"w": self.w.cpu().numpy(), "b": self.b.cpu().numpy()}


class UNet:
    def __init__(self, channels=CHANNELS, seed=0):
        gen = np.random.default_rng(seed)
        self.ch = tuple(channels)
        L = len(self.ch)
        self.enc = []
        cin = IN_CHANNELS
        for l in range(L):
            self.enc.append([Conv(cin, self.ch[l], True, gen), Conv(self.ch[l], self.ch[l], True, gen)])
            cin = self.ch[l]
        self.dec = [None] * (L - 1)
        for l in range(L - 2, -1, -1):
            self.dec[l] = [Conv(self.ch[l + 1] + self.ch[l], self.ch[l], True, gen), Conv(self.ch[l], self.ch[l], True, gen)]
        self.out = Conv(self.ch[0], OUT_CHANNELS, False, gen)
        self.step_count = 0

    @property
    def multiple(self):
        return 1 << (len(self.ch) - 1)

    def layers(self):
        """convs in forward order: encoder levels 0..L-1, decoder levels L-2..0, output"""
        out = []
        for a in self.enc:
            out += a
        for a in reversed(self.dec):
            out += a
        return out + [self.out]

    def forward(self, x, keep=False):
        lib = load()
        s = _stream()
        L = len(self.ch)
        N, H, W, _ = x.shape
        assert H % self.multiple == 0 and W % self.multiple == 0, "H and W must be multiples of %d" % self.multiple
        self.skips = []
        h = x
        pooled = None
        for l in range(L):
            if l > 0:
                h = pooled   # the 2x2 average of the previous level's output, written by its last conv
            h = self.enc[l][0].forward(h, keep)
            if l < L - 1:
                h, pooled = self.enc[l][1].forward_pool(h, keep)
            else:
                h = self.enc[l][1].forward(h, keep)
            self.skips.append(h)
        for l in range(L - 2, -1, -1):
            skip = self.skips[l]
            if keep:    # training keeps the concatenation: the weight gradient reads it
                n_, hh, ww, c2 = skip.shape
                c1 = h.shape[3]
                cat = bf16_empty((n_, hh, ww, c1 + c2))
                check(lib.ydn_up2_concat(_p(h), c1, _p(skip), c2, n_, hh, ww, _p(cat), s))
                h = self.dec[l][0].forward(cat, keep)
            else:
                h = self.dec[l][0].forward_up2cat(h, skip)
            h = self.dec[l][1].forward(h, keep)
        y = self.out.forward(h, keep)
        if not keep:
            self.skips = []
        return y

    def backward(self, g):
        lib = load()
        s = _stream()
        L = len(self.ch)
        g = self.out.backward(g)
        gskip = [None] * L
        for l in range(L - 1):
            g = self.dec[l][1].backward(g)
            g = self.dec[l][0].backward(g)
            n_, hh, ww, _ = g.shape
            c1, c2 = self.ch[l + 1], self.ch[l]
            gin = bf16_empty((n_, hh // 2, ww // 2, c1))
            gs = bf16_empty((n_, hh, ww, c2))
            check(lib.ydn_up2_concat_grad(_p(g), c1, c2, n_, hh, ww, _p(gin), _p(gs), 0, s))
            gskip[l] = gs
            g = gin
        for l in range(L - 1, -1, -1):
            if l < L - 1:
                check(lib.ydn_add(_p(g), _p(gskip[l]), g.numel(), s))
            g = self.enc[l][1].backward(g)
            g = self.enc[l][0].backward(g, need_input_grad=l > 0)
            if l > 0:
                n_, hh, ww, c = g.shape
                gp = bf16_empty((n_, hh * 2, ww * 2, c))
                check(lib.ydn_pool2_grad(_p(g), n_, hh * 2, ww * 2, c, _p(gp), s))
                g = gp
        self.skips = []

    def zero_grad(self):
        for c in self.layers():
            c.zero_grad()

    def step(self, lr):
        self.step_count += 1
        for c in self.layers():
            c.step(lr, self.step_count)

    def state(self):
        return {"channels": self.ch, "layers": [c.params() for c in self.layers()]}

    def load_state(self, st):
        import torch
        for c, p in zip(self.layers(), st["layers"]):
            c.w.copy_(torch.from_numpy(np.ascontiguousarray(p["w"])).cuda())
            c.b.copy_(torch.from_numpy(np.ascontiguousarray(p["b"])).cuda())
            c.w16 = to_bf16_bits(c.w)


# ---------------------------------------------------------------- data and training
def pack_input(noisy, features, pad_to=16):
    """noisy: (N, H, W, 4) float32 cuda, features: (N, H, W, 12) float32 cuda -> (N, H', W, 16) bf16 with
    H' = H rounded up to `pad_to` (zero rows appended)."""
    N, H, W, _ = noisy.shape
    Hp = (H + pad_to - 1) // pad_to * pad_to
    x = bf16_empty((N, Hp, W, IN_CHANNELS))
    x.zero_()
    for n in range(N):
        check(load().ydn_pack_input(_p(noisy[n]), _p(features[n]), H * W, IN_CHANNELS, _p(x[n]), _stream()))
    return x


def loss_and_grad(net_out, noisy, ref, want_grad=True, want_pred=False):
    """L1 in log space over the first H rows (the rest is padding)."""
    import torch
    N, Hp, W, C = net_out.shape
    H = noisy.shape[1]
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    g = bf16_empty(net_out.shape) if want_grad else None
    if g is not None:
        g.zero_()
    pred = torch.empty((N, H, W, 4), dtype=torch.float32, device="cuda") if want_pred else None
    for n in range(N):
        check(load().ydn_l1_log_loss(_p(net_out[n]), C, _p(noisy[n]), _p(ref[n]), H * W, _p(g[n]) if g is not None else None,
                                     _p(loss), _p(pred[n]) if pred is not None else None, _stream()))
    return loss / N, g, pred


def denoise(net, noisy, features):
    """(H, W, 4) noisy film + (H, W, 12) features (cuda float32) -> (H, W, 4) denoised film."""
    x = pack_input(noisy[None], features[None], net.multiple)
    y = net.forward(x, keep=False)
    _, _, pred = loss_and_grad(y, noisy[None], noisy[None], want_grad=False, want_pred=True)
    return pred[0]


def rel_mse(img, ref, eps=1e-2):
    """mean over pixels and channels of (img - ref)^2 / (ref^2 + eps)"""
    a = np.asarray(img, np.float64)[..., :3]
    r = np.asarray(ref, np.float64)[..., :3]
    return float(np.mean((a - r) ** 2 / (r ** 2 + eps)))


def render_pair(it, frame, spp_in, spp_ref, start_ref):
    """noisy film + first-hit features of passes [0, spp_in) and a reference of [start_ref, +spp_ref)"""
    it.set_frame(frame)
    it.clear_film()
    it.render(0, spp_in)
    noisy = it.film()
    feats = it.first_hit_features()
    it.clear_film()
    it.render(start_ref, spp_ref)
    ref = it.film()
    return noisy, feats, ref


def _training_params(rng, n_scenes):
    """the draws of each training scene, in scene order: shapes (100 to 600), quad lights (1 to 4),
    emissive triangles (0 to 5), camera eye and target jitter and field of view"""
    out = []
    for _ in range(n_scenes):
        n_shapes = int(rng.integers(100, 600))
        n_lights = int(rng.integers(1, 5))
        n_emissive = int(rng.integers(0, 6))
        out.append((n_shapes, n_lights, n_emissive, rng.normal(0, 3.0, 3), rng.normal(0, 2.0, 3), float(rng.uniform(0.6, 1.2))))
    return out


def _training_scene(params, seed, k):
    from . import camera, scenes
    n_shapes, n_lights, n_emissive, deye, dtgt, fovy = params
    scene, cam = scenes.random_shapes(n_shapes=n_shapes, seed=int(seed * 1000 + k), cube=12.0, n_lights=n_lights,
                                      emissive_triangles=n_emissive)
    eye = np.array(cam.eye, np.float64) + deye
    tgt = np.array(cam.target, np.float64) + dtgt
    return scene, camera.Camera(eye=tuple(eye), target=tuple(tgt), fovy=fovy)


def make_dataset(n_scenes, res=256, spp_in=4, spp_ref=1024, seed=100, log=None, workers=3, prefetch=6):
    """Renders of seeded random-shapes scenes with the NEE integrator: noisy inputs, features and
    references, as float32 arrays. Worker threads generate the next scenes and build their BVHs
    (resident scenes) while the GPU renders the current one; every scene's draws are made up front
    from one seeded generator, so the dataset does not depend on the timing."""
    import threading
    from . import _lib, camera
    params = _training_params(np.random.default_rng(seed), n_scenes)
    ready = {}
    cv = threading.Condition()
    state = {"next": 0, "consumed": 0, "error": None}

    def producer():
        while True:
            with cv:
                while state["next"] < n_scenes and state["next"] >= state["consumed"] + prefetch and state["error"] is None:
                    cv.wait()
                if state["next"] >= n_scenes or state["error"] is not None:
                    return
                k = state["next"]
                state["next"] += 1
            try:
                scene, cam = _training_scene(params[k], seed, k)
                item = (_lib.ResidentScene(scene), cam, params[k][0])
            except Exception as e:  # reported by the consumer
                with cv:
                    state["error"] = e
                    cv.notify_all()
                return
            with cv:
                ready[k] = item
                cv.notify_all()

    threads = [threading.Thread(target=producer, daemon=True) for _ in range(max(1, workers))]
    for t in threads:
        t.start()
    it = _lib.Integrator(_lib.YPT_INTEGRATOR_NEE, 0)
    it.set_option("first_hit_features", 1)
    data = []
    for k in range(n_scenes):
        with cv:
            while k not in ready and state["error"] is None:
                cv.wait()
            if k not in ready:
                raise state["error"]
            rs, cam, n_shapes = ready.pop(k)
        it.set_scene(rs)
        fr = camera.make_frame(cam, res, res, max_depth=8)
        noisy, feats, ref = render_pair(it, fr, spp_in, spp_ref, 1000)
        data.append((noisy, feats, ref))
        rs.close()
        with cv:
            state["consumed"] = k + 1
            cv.notify_all()
        if log and (k % 50 == 0 or k == n_scenes - 1):
            log("scene %d/%d: %d shapes" % (k + 1, n_scenes, n_shapes))
    for t in threads:
        t.join()
    it.close()
    return data


def train(net, data, iters=2000, batch=8, crop=128, lr=1e-3, seed=0, log=None, log_every=100):
    """Adam on random crops and flips of the dataset; returns the loss history."""
    import torch
    rng = np.random.default_rng(seed)
    noisy = torch.from_numpy(np.stack([d[0] for d in data])).cuda()
    feats = torch.from_numpy(np.stack([d[1] for d in data])).cuda()
    refs = torch.from_numpy(np.stack([d[2] for d in data])).cuda()
    S, H, W, _ = noisy.shape
    hist = []
    for i in range(iters):
        idx = rng.integers(0, S, size=batch)
        ys = rng.integers(0, H - crop + 1, size=batch)
        xs = rng.integers(0, W - crop + 1, size=batch)
        nb = torch.stack([noisy[a, y:y + crop, x:x + crop] for a, y, x in zip(idx, ys, xs)]).contiguous()
        fb = torch.stack([feats[a, y:y + crop, x:x + crop] for a, y, x in zip(idx, ys, xs)]).contiguous()
        rb = torch.stack([refs[a, y:y + crop, x:x + crop] for a, y, x in zip(idx, ys, xs)]).contiguous()
        x = pack_input(nb, fb, net.multiple)
        net.zero_grad()
        y = net.forward(x, keep=True)
        loss, g, _ = loss_and_grad(y, nb, rb)
        net.backward(g)
        step_lr = lr * 0.5 * (1 + math.cos(math.pi * i / iters))
        net.step(step_lr)
        if i % log_every == 0 or i == iters - 1:
            l = float(loss.item())
            hist.append((i, l))
            if log:
                log("iter %d loss %.5f" % (i, l))
    return hist
