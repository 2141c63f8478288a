"""Benchmark of the spectral path-tracing core on one MI355X.

Workload (default): next-event-estimation path tracing of the ~1M-triangle random-shapes scene at
1920x1080, path depth 12, one Sobol sample index per step (a "step" is one full-image pass of the hot
path: camera rays, extension/shadow traversal, shading, film merge). Prints one JSON line with
path samples/s, the roofline of the dominant kernel (BVH extension traversal, HBM-bound) and the
CPU checker's throughput on a bounded sample of the same workload.

--integrator bdpt: the bidirectional workload (the same scene plus 8 seeded blackbody emissive
triangles, 1280x720, camera subpaths of depth 8, light subpaths of 9 vertices, i.e. depth 8); its
dominant kernel is the camera-subpath kernel.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NODE_BYTES = {2: 64, 8: 256}
TRI_BYTES = 48
RAY_BYTES = 4 + 32 + 16          # queue entry + origin/direction record + hit record
SHADOW_RAY_BYTES = 4 + 48 + 16   # queue entry + shadow record + light-triangle vertices (3x12 rounded)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=64)
    p.add_argument("--warmup", type=int, default=16)
    p.add_argument("--width", type=int, default=0, help="0: 1920 (nee, bpt) or 1280 (bdpt)")
    p.add_argument("--height", type=int, default=0, help="0: 1080 (nee, bpt) or 720 (bdpt)")
    p.add_argument("--depth", type=int, default=0, help="0: 12 (nee, bpt) or 8 (bdpt)")
    p.add_argument("--shapes", type=int, default=2000)
    p.add_argument("--integrator", default="nee", choices=["nee", "bpt", "bdpt"])
    p.add_argument("--bvh-width", type=int, default=0, help="2 or 8; 0 = autotuned choice")
    p.add_argument("--block-size", type=int, default=0)
    p.add_argument("--lds-depth", type=int, default=0)
    p.add_argument("--passes-in-flight", type=int, default=0, help="sample passes per wavefront batch; 0 = tuned")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU work of the checker sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--spectral-extensions", action="store_true",
                   help="BK7 Sellmeier glass and Planck-spectrum emitters instead of the reference's Cauchy and RGB x D65")
    p.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                   help="launch option over the tuned choice (repeatable)")
    return p.parse_args()


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: the path tracer runs on a single MI355X in this build "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from yapt_amd import _lib, camera, scenes, tuning

    bdpt = a.integrator == "bdpt"
    integ = {"nee": _lib.YPT_INTEGRATOR_NEE, "bpt": _lib.YPT_INTEGRATOR_BPT, "bdpt": _lib.YPT_INTEGRATOR_BDPT}[a.integrator]
    a.width = a.width or (1280 if bdpt else 1920)
    a.height = a.height or (720 if bdpt else 1080)
    a.depth = a.depth or (8 if bdpt else 12)
    light_vertices = 9 if bdpt else 16
    t0 = time.time()
    scene, cam = scenes.random_shapes(n_shapes=a.shapes, seed=2, emissive_triangles=8 if bdpt else 0,
                                      spectral_extensions=a.spectral_extensions)
    frame = camera.make_frame(cam, a.width, a.height, max_depth=a.depth, light_path_vertices=light_vertices)
    t_scene = time.time() - t0

    it = _lib.Integrator(integ, 0)
    cfg = tuning.best_config("%s-shapes%d-%dx%d-d%d" % (a.integrator, a.shapes, a.width, a.height, a.depth))
    if a.bvh_width:
        cfg["bvh_width"] = a.bvh_width
    if a.block_size:
        cfg["block_size"] = a.block_size
    if a.lds_depth:
        cfg["lds_stack_depth"] = a.lds_depth
    if a.passes_in_flight:
        cfg["passes_in_flight"] = a.passes_in_flight
    for o in a.option:
        k, v = o.split("=", 1)
        cfg[k.strip()] = int(v)
    for k, v in cfg.items():
        it.set_option(k, v)
    width = cfg["bvh_width"]
    t1 = time.time()
    it.set_scene(scene)
    it.set_frame(frame)
    t_upload = time.time() - t1

    # warmup
    it.render(0, a.warmup)
    it.synchronize()

    # timed region: K passes, inputs resident in HBM
    it.synchronize()
    ts = time.perf_counter()
    it.render(a.warmup, a.steps)
    it.synchronize()
    te = time.perf_counter()
    secs = te - ts
    samples = a.width * a.height * a.steps
    value = samples / secs

    # kernel timing of the same passes (events on the render stream) and algorithmic work counts
    it.set_option("profile", 1)
    it.clear_film()
    it.render(a.warmup, a.steps)
    st_prof = it.stats()
    it.set_option("profile", 0)
    it.set_option("count_traversal", 1)
    it.clear_film()
    it.render(a.warmup, a.steps)
    st_cnt = it.stats()
    it.set_option("count_traversal", 0)
    launches = st_prof[10]
    ext_ms = st_prof[2]
    if bdpt:   # camera-subpath kernel: traversal bytes of its extension and connection rays
        ext_bytes = st_cnt[6] * NODE_BYTES[width] + st_cnt[7] * TRI_BYTES
    else:
        ext_bytes = st_cnt[6] * NODE_BYTES[width] + st_cnt[7] * TRI_BYTES + st_cnt[0] * RAY_BYTES
    ext_bytes_per_launch = ext_bytes / launches
    ext_ms_per_launch = ext_ms / launches
    achieved = ext_bytes_per_launch / (ext_ms_per_launch * 1e-3) / 1e9
    traffic = tuning.measured_traffic(("camera-%s" if bdpt else "extend-%s") % a.integrator)

    # CPU checker on a bounded sample of the same workload
    cpu = None
    if not a.no_cpu:
        from oracle import oracle_lib
        threads = cpu_threads()
        stride = 64
        probe = oracle_lib.render(scene, frame, integ, a.warmup, 1, threads=threads, use_bvh=True, pixel_stride=stride)
        per_sample = probe["stats"][0] / max(1.0, probe["stats"][1])
        target = a.cpu_seconds
        want = max(1.0, target / max(per_sample, 1e-9))
        npix = a.width * a.height
        stride = int(max(1, min(4096, npix / want)))
        passes = int(max(1, min(a.steps, round(want / npix)))) if stride == 1 else 1
        res = oracle_lib.render(scene, frame, integ, a.warmup, passes, threads=threads, use_bvh=True, pixel_stride=stride)
        cpu = {"value": res["stats"][1] / res["stats"][0], "unit": "path samples/s", "cores": threads, "kind": "port",
               "sample": "every %d-th %s of %d %dx%d pass(es) from sample index %d, %d paths, %.1f s on %d threads of %s" %
                         (stride, "8x8 camera workgroup" if bdpt else "pixel", passes, a.width, a.height, a.warmup,
                          int(res["stats"][1]), res["stats"][0], threads, cpu_model())}

    if bdpt:
        metric = "path samples/s (BDPT, random shapes ~1M tris + 8 emissive triangles, 1280x720, depth 8)"
        kernel = ("camera-subpath stage: k_extend_p<%d> + k_bdpt_cshade + k_shadow_p<%d> (extension and connection rays)" % (width, width)
                  if cfg.get("bdpt_wavefront", 1) else "k_bdpt_camera<%d> (camera subpaths with all connections)" % width)
        kernel_ms = {"camera": st_prof[2], "light": st_prof[3], "splat": st_prof[4], "headers": st_prof[11],
                     "sort_sum_merge": st_prof[12], "render_total": st_prof[5]}
        rays = {"camera_side": st_cnt[0], "light_side": st_cnt[1]}
        trav = {"camera_nodes": st_cnt[6], "camera_tris": st_cnt[7], "light_nodes": st_cnt[8], "light_tris": st_cnt[9]}
    else:
        metric = "path samples/s (NEE, random shapes ~1M tris, 1920x1080, depth 12)"
        kernel = "%s<%d> (closest-hit traversal)" % ("k_extend_p" if cfg.get("persistent", 1) else "k_extend", width)
        kernel_ms = {"extend": ext_ms, "shade": st_prof[4], "shadow": st_prof[3], "start": st_prof[11],
                     "merge": st_prof[12], "render_total": st_prof[5]}
        rays = {"extension": st_cnt[0], "shadow": st_cnt[1]}
        trav = {"ext_nodes": st_cnt[6], "ext_tris": st_cnt[7], "shadow_nodes": st_cnt[8], "shadow_tris": st_cnt[9]}
    out = {
        "metric": metric,
        "value": value,
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": secs * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-shapes scene, seed 2)",
        "config": {"workload": "%s random-shapes %d shapes / %d tris, %dx%d, depth %d, 1 spp per step" %
                               (a.integrator.upper(), a.shapes, scene.triangle_count, a.width, a.height, a.depth),
                   "launch": dict(cfg, passes_in_flight=int(st_prof[13]))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
                     "bytes_per_launch": ext_bytes_per_launch, "ms_per_launch": ext_ms_per_launch},
        "cpu_baseline": cpu,
        "detail": {
            "scene_build_s": t_scene, "upload_and_bvh_s": t_upload,
            "kernel_ms": kernel_ms, "rays": rays, "traversal": trav,
            "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        },
    }
    print(json.dumps(out))
    it.close()


if __name__ == "__main__":
    main()
