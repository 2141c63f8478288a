"""Integrator agreement in the reference's IntegratorCompare style (compare_integrators.py, report.py):
render one scene with BPT, BPT_WITH_NEE and BDPT, take each FinalColor capture (ACES, manual
exposure, 8-bit sRGB), and print the pairwise difference figures. "rel|d| 16x" is the relative mean
absolute difference after a box downsample, which removes Monte Carlo noise and leaves disagreement.

    python -m yapt_amd.compare [--scene punctual] [--light-type mixed] [--seconds 25 | --spp N]
                               [--width 960 --height 704] [--out DIR]

Prints one JSON object: per-pair rel_full, rel_down (the 16x figure) and mean_gap in percent.
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


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--scene", default="punctual", choices=["punctual", "cornell"])
    p.add_argument("--light-type", default=os.environ.get("LIGHT_TYPE", "mixed"))
    p.add_argument("--seconds", type=float, default=float(os.environ.get("ACCUM_S", "25")))
    p.add_argument("--spp", type=int, default=0)
    p.add_argument("--width", type=int, default=960)
    p.add_argument("--height", type=int, default=704)
    p.add_argument("--depth", type=int, default=8)
    p.add_argument("--out", default=None, help="folder for the PNG captures")
    a = p.parse_args()
    from . import _lib, camera, scenes
    if a.scene == "punctual":
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
                      "resolution": [a.width, a.height], "spp": spp, "pairs": pairs}))


if __name__ == "__main__":
    main()
