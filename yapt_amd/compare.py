"""Integrator agreement in the reference's IntegratorCompare style (compare_integrators.py, report.py,
focus_sweep.py): render one scene with BPT, BPT_WITH_NEE and BDPT, take each FinalColor capture (ACES,
manual exposure, 8-bit sRGB), and print the pairwise difference figures. "rel|d| 16x" is the relative
mean absolute difference after a box downsample, which removes Monte Carlo noise and leaves
disagreement.

    python -m yapt_amd.compare [--scene focused|punctual|cornell] [--focus 8] [--light-type mixed]
                               [--seconds 25 | --spp N] [--width 900 --height 700] [--out DIR]
    python -m yapt_amd.compare --sweep 0,2,8,32 [--integrator 1]     # focus_sweep.py

The path depth defaults to the reference's maxRayDepth of 32. Prints one JSON object: per-pair
rel_full, rel_down (the 16x figure) and mean_gap in percent; with --sweep, the capture mean, p99 and the
floor profile (row 78% down the image, 12 buckets) per focus.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

INTEGRATORS = [(0, "bpt"), (1, "bpt_nee"), (2, "bdpt")]


def difference_stats(a, b, downsample=(15, 11)):
    """report.py:59-85 difference figures of two 8-bit captures"""
    from PIL import Image
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    mid = (a.mean() + b.mean()) / 2
    sa = np.asarray(Image.fromarray(a.astype(np.uint8)).resize(downsample, Image.BOX), np.float64)
    sb = np.asarray(Image.fromarray(b.astype(np.uint8)).resize(downsample, Image.BOX), np.float64)
    smid = (sa.mean() + sb.mean()) / 2
    return {"mean_a": float(a.mean()), "mean_b": float(b.mean()), "mean_abs": float(d.mean()),
            "rel_full": float(100.0 * d.mean() / mid) if mid else 0.0,
            "rel_down": float(100.0 * np.abs(sa - sb).mean() / smid) if smid else 0.0,
            "mean_gap": float(100.0 * (b.mean() - a.mean()) / a.mean()) if a.mean() else 0.0,
            "p99": float(np.percentile(d, 99)), "max": float(d.max())}


def capture(lib, scene, frame, integrator, seconds=None, spp=None, chunk=16):
    """accumulates for a time budget (or a fixed spp) and returns the RGB8 FinalColor and the spp"""
    it = lib.Integrator(integrator, 0)
    it.set_scene(scene)
    it.set_frame(frame)
    n = 0
    t0 = time.perf_counter()
    while True:
        k = chunk if spp is None else min(chunk, spp - n)
        if k <= 0:
            break
        it.render(n, k)
        n += k
        if spp is None:
            it.synchronize()
            if time.perf_counter() - t0 >= seconds:
                break
    rgba = it.final_color()
    it.close()
    return rgba[..., :3], n


def floor_profile(image):
    """focus_sweep.py print_profile: luma (PIL "L") of the row 78% down the image, in 12 buckets"""
    from PIL import Image
    luma = np.asarray(Image.fromarray(np.asarray(image, np.uint8)).convert("L"), np.float64)
    row = luma[int(luma.shape[0] * 0.78)]
    return [float(v) for v in row[: len(row) // 12 * 12].reshape(12, -1).mean(axis=1)]


def sweep(a):
    """focus_sweep.py: one integrator over the emission focus values of the focused-emitter box"""
    from . import _lib, camera, scenes
    out = {}
    for focus in [float(v) for v in a.sweep.split(",")]:
        scene, cam = scenes.focused_emission(focus, two_sided=a.two_sided)
        frame = camera.make_frame(cam, a.width, a.height, max_depth=a.depth)
        img, n = capture(_lib, scene, frame, a.integrator, seconds=a.seconds, spp=a.spp or None)
        out["%g" % focus] = {"spp": n, "mean": float(img.mean()), "p99": float(np.percentile(img, 99)),
                             "floor_profile": floor_profile(img)}
        print("  focus %g: %d spp, mean %.2f" % (focus, n, img.mean()), file=sys.stderr, flush=True)
    print(json.dumps({"scene": "focused", "integrator": a.integrator, "two_sided": a.two_sided,
                      "resolution": [a.width, a.height], "depth": a.depth, "sweep": out}))


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--scene", default="focused", choices=["focused", "punctual", "cornell"])
    p.add_argument("--focus", type=float, default=float(os.environ.get("EMISSION_FOCUS", "8")))
    p.add_argument("--two-sided", action="store_true", default=os.environ.get("TWO_SIDED", "0") == "1")
    p.add_argument("--light-type", default=os.environ.get("LIGHT_TYPE", "mixed"))
    p.add_argument("--seconds", type=float, default=float(os.environ.get("ACCUM_S", "25")))
    p.add_argument("--spp", type=int, default=0)
    p.add_argument("--width", type=int, default=900)
    p.add_argument("--height", type=int, default=700)
    p.add_argument("--depth", type=int, default=32)
    p.add_argument("--sweep", default=None, help="comma separated emission focus values (focus_sweep.py)")
    p.add_argument("--integrator", type=int, default=int(os.environ.get("INTEGRATOR", "1")), help="for --sweep")
    p.add_argument("--out", default=None, help="folder for the PNG captures")
    a = p.parse_args()
    if a.sweep:
        return sweep(a)
    from . import _lib, camera, scenes
    if a.scene == "focused":
        (scene, cam), env = scenes.focused_emission(a.focus, two_sided=a.two_sided), 0.0
    elif a.scene == "punctual":
        scene, cam, env = scenes.punctual_lights(a.light_type)
    else:
        (scene, cam), env = scenes.cornell_box(), 0.0
    frame = camera.make_frame(cam, a.width, a.height, max_depth=a.depth, env_intensity_scale=env)
    caps, spp = {}, {}
    for integ, label in INTEGRATORS:
        caps[label], spp[label] = capture(_lib, scene, frame, integ, seconds=a.seconds, spp=a.spp or None)
        print("  %s: %d spp" % (label, spp[label]), file=sys.stderr, flush=True)
    if a.out:
        from PIL import Image
        os.makedirs(a.out, exist_ok=True)
        for label, img in caps.items():
            Image.fromarray(img).save(os.path.join(a.out, label + ".png"))
    pairs = {"%s vs %s" % (x, y): difference_stats(caps[x], caps[y]) for x, y in itertools.combinations(caps, 2)}
    print(json.dumps({"scene": a.scene, "light_type": a.light_type if a.scene == "punctual" else None,
                      "focus": a.focus if a.scene == "focused" else None, "resolution": [a.width, a.height],
                      "depth": a.depth, "spp": spp, "pairs": pairs}))


if __name__ == "__main__":
    main()
