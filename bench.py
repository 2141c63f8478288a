"""Benchmark of the spectral path-tracing core on one MI355X.

Workload (default): next-event-estimation path tracing of the ~1M-triangle random-shapes scene at
1920x1080, path depth 12, one Sobol sample index per step (a "step" is one full-image pass of the hot
path: camera rays, extension/shadow traversal, shading, film merge). Prints one JSON line with
path samples/s, the roofline of the dominant kernel (BVH extension traversal, priced against HBM) and
the CPU checker's throughput on a bounded sample of the same workload.

Under "detail", unless --no-extras: the bidirectional workload (the same scene plus 8 seeded blackbody
emissive triangles, 1280x720, camera and light subpaths of depth 8), a fixed-length serving trace
(Poisson arrivals at 40 jobs/s over resident scenes of ~10k, 100k, 1M and 5M triangles) with its
p50/p99 latency, and the learned denoiser's judgement at 1920x1080 (4 spp + denoiser against plain
64 spp, relMSE against a plain reference render, wall time of each).

--integrator bdpt: the bidirectional workload as the headline; its dominant kernel is the
camera-subpath stage.
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
CPU_SHARE = 16                   # host threads per GPU on the benchmark hosts


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
    p.add_argument("--no-extras", action="store_true", help="skip the BDPT, serving and denoiser entries of detail")
    p.add_argument("--spectral-extensions", action="store_true",
                   help="BK7 Sellmeier glass and Planck-spectrum emitters instead of the reference's Cauchy and RGB x D65")
    p.add_argument("--option", action="append", default=[], metavar="NAME=VALUE",
                   help="launch option over the tuned choice (repeatable)")
    return p.parse_args()


def cpu_threads():
    """the host threads of the checker: every core the process may use, up to the CPU share of one GPU"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(CPU_SHARE, n)), n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_workload(integrator, shapes, width, height, depth, steps, warmup, overrides=(), spectral=False):
    """Renders `warmup` untimed then `steps` timed passes of one workload (inputs resident in HBM), then
    the same passes with kernel timing (HIP events on the render stream) and with traversal counting.
    Returns a dict with the rate, the launch configuration, per-kernel times and traversal counts."""
    from yapt_amd import _lib, camera, scenes, tuning
    bdpt = integrator == "bdpt"
    integ = {"nee": _lib.YPT_INTEGRATOR_NEE, "bpt": _lib.YPT_INTEGRATOR_BPT, "bdpt": _lib.YPT_INTEGRATOR_BDPT}[integrator]
    light_vertices = 9 if bdpt else 16
    t0 = time.time()
    scene, cam = scenes.random_shapes(n_shapes=shapes, seed=2, emissive_triangles=8 if bdpt else 0,
                                      spectral_extensions=spectral)
    frame = camera.make_frame(cam, width, height, max_depth=depth, light_path_vertices=light_vertices)
    t_scene = time.time() - t0
    it = _lib.Integrator(integ, 0)
    cfg = tuning.best_config("%s-shapes%d-%dx%d-d%d" % (integrator, shapes, width, height, depth))
    for k, v in overrides:
        cfg[k] = v
    for k, v in cfg.items():
        it.set_option(k, v)
    t1 = time.time()
    it.set_scene(scene)
    it.set_frame(frame)
    t_upload = time.time() - t1

    it.render(0, warmup)
    it.synchronize()
    # timed region: `steps` passes, inputs resident in HBM
    ts = time.perf_counter()
    it.render(warmup, steps)
    it.synchronize()
    secs = time.perf_counter() - ts

    it.set_option("profile", 1)
    it.clear_film()
    it.render(warmup, steps)
    st_prof = it.stats()
    it.set_option("profile", 0)
    it.set_option("count_traversal", 1)
    it.clear_film()
    it.render(warmup, steps)
    st_cnt = it.stats()
    it.set_option("count_traversal", 0)
    it.close()
    return {"integ": integ, "scene": scene, "frame": frame, "cfg": cfg, "secs": secs, "steps": steps,
            "value": width * height * steps / secs, "prof": st_prof, "cnt": st_cnt,
            "t_scene": t_scene, "t_upload": t_upload}


def dominant_kernel(w, integrator):
    """algorithmic bytes and time per launch of the dominant kernel (see DESIGN.md for the byte counts)"""
    from yapt_amd import tuning
    bdpt = integrator == "bdpt"
    width = w["cfg"]["bvh_width"]
    st_prof, st_cnt = w["prof"], w["cnt"]
    launches = st_prof[10]
    if bdpt:   # camera-subpath stage: traversal bytes of its extension and connection rays
        nbytes = st_cnt[6] * NODE_BYTES[width] + st_cnt[7] * TRI_BYTES
        kernel = ("camera-subpath stage: k_extend_p<%d> + k_bdpt_cshade + k_bdpt_connect + k_shadow_p<%d> (extension and "
                  "connection rays)" % (width, width) if w["cfg"].get("bdpt_wavefront", 1) else "k_bdpt_camera<%d>" % width)
        key = "camera-bdpt"
    else:
        nbytes = st_cnt[6] * NODE_BYTES[width] + st_cnt[7] * TRI_BYTES + st_cnt[0] * RAY_BYTES
        kernel = "%s<%d> (closest-hit traversal)" % ("k_extend_p" if w["cfg"].get("persistent", 1) else "k_extend", width)
        key = "extend-%s" % integrator
    per_launch = nbytes / launches
    ms = st_prof[2] / launches
    achieved = per_launch / (ms * 1e-3) / 1e9
    traffic = tuning.measured_traffic(key)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic.get("bytes_per_launch") if traffic else None, "kernel": kernel,
            "bytes_per_launch": per_launch, "ms_per_launch": ms}
    if traffic:
        share = traffic["bytes_per_launch"] / per_launch
        roof["traffic_profiled_at"] = traffic.get("launch")
        roof["limiter"] = ("latency and the L2/MALL hierarchy, not HBM bandwidth: measured HBM traffic is %.0f%% of "
                           "the algorithmic bytes (%.2f of %.2f GB per launch), %.2f TB/s at this launch time"
                           % (100 * share, traffic["bytes_per_launch"] / 1e9, per_launch / 1e9,
                              traffic["bytes_per_launch"] / (ms * 1e-3) / 1e12)) if share < 0.5 else "HBM bandwidth"
    return roof


def cpu_baseline(w, integrator, warmup, steps, cpu_seconds):
    """the oracle (a scalar C++ port, std::thread over pixel tiles) on a bounded sample of the workload"""
    from oracle import oracle_lib
    threads, ncores = cpu_threads()
    scene, frame = w["scene"], w["frame"]
    probe = oracle_lib.render(scene, frame, w["integ"], warmup, 1, threads=threads, use_bvh=True, pixel_stride=64)
    per_sample = probe["stats"][0] / max(1.0, probe["stats"][1])
    want = max(1.0, cpu_seconds / max(per_sample, 1e-9))
    npix = frame.width * frame.height
    stride = int(max(1, min(4096, npix / want)))
    passes = int(max(1, min(steps, round(want / npix)))) if stride == 1 else 1
    res = oracle_lib.render(scene, frame, w["integ"], warmup, passes, threads=threads, use_bvh=True, pixel_stride=stride)
    return {"value": res["stats"][1] / res["stats"][0], "unit": "path samples/s", "cores": threads, "kind": "port",
            "sample": "every %d-th %s of %d %dx%d pass(es) from sample index %d, %d paths, %.1f s on %d threads "
                      "(the CPU share of one GPU; the host shows %d cores) of %s; binned-SAH BVH, near child first"
                      % (stride, "8x8 camera workgroup" if integrator == "bdpt" else "pixel", passes, frame.width,
                         frame.height, warmup, int(res["stats"][1]), res["stats"][0], threads, ncores, cpu_model())}


def kernel_detail(w, integrator):
    st_prof, st_cnt = w["prof"], w["cnt"]
    if integrator == "bdpt":
        return ({"camera": st_prof[2], "light": st_prof[3], "splat": st_prof[4], "headers": st_prof[11],
                 "sort_sum_merge": st_prof[12], "render_total": st_prof[5]},
                {"camera_side": st_cnt[0], "light_side": st_cnt[1]},
                {"camera_nodes": st_cnt[6], "camera_tris": st_cnt[7], "light_nodes": st_cnt[8], "light_tris": st_cnt[9]})
    return ({"extend": st_prof[2], "shade": st_prof[4], "shadow": st_prof[3], "start": st_prof[11],
             "merge": st_prof[12], "render_total": st_prof[5]},
            {"extension": st_cnt[0], "shadow": st_cnt[1]},
            {"ext_nodes": st_cnt[6], "ext_tris": st_cnt[7], "shadow_nodes": st_cnt[8], "shadow_tris": st_cnt[9]})


def extras(a):
    out = {}
    t = time.time()
    log("detail: BDPT workload")
    w = run_workload("bdpt", a.shapes, 1280, 720, 8, 16, 4)
    kms, rays, _ = kernel_detail(w, "bdpt")
    out["bdpt"] = {"workload": "BDPT random-shapes %d shapes + 8 blackbody emissive triangles / %d tris, 1280x720, "
                               "camera and light subpaths of depth 8" % (a.shapes, w["scene"].triangle_count),
                   "value": w["value"], "unit": "path samples/s", "ms_per_pass": w["secs"] * 1e3 / w["steps"],
                   "passes": w["steps"], "roofline": dominant_kernel(w, "bdpt"), "kernel_ms": kms,
                   "launch": w["cfg"], "wall_s": time.time() - t}
    del w
    t = time.time()
    log("detail: serving trace")
    from yapt_amd import serving
    srv = serving.trace(jobs=48, rate=40.0, log=log)
    srv["wall_s"] = time.time() - t
    out["serving"] = srv
    t = time.time()
    log("detail: denoiser judgement")
    from yapt_amd import denoise_bench
    den = denoise_bench.judge(denoise_bench.load_weights(), 1920, 1080, ref_spp=1024, log=log)
    den["wall_s_total"] = time.time() - t
    out["denoiser"] = den
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU: the path tracer runs on a single MI355X in this build "
              "(multi-GPU sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    bdpt = a.integrator == "bdpt"
    a.width = a.width or (1280 if bdpt else 1920)
    a.height = a.height or (720 if bdpt else 1080)
    a.depth = a.depth or (8 if bdpt else 12)
    over = []
    if a.bvh_width:
        over.append(("bvh_width", a.bvh_width))
    if a.block_size:
        over.append(("block_size", a.block_size))
    if a.lds_depth:
        over.append(("lds_stack_depth", a.lds_depth))
    if a.passes_in_flight:
        over.append(("passes_in_flight", a.passes_in_flight))
    for o in a.option:
        k, v = o.split("=", 1)
        over.append((k.strip(), int(v)))
    w = run_workload(a.integrator, a.shapes, a.width, a.height, a.depth, a.steps, a.warmup, over, a.spectral_extensions)
    roof = dominant_kernel(w, a.integrator)
    cpu = None if a.no_cpu else cpu_baseline(w, a.integrator, a.warmup, a.steps, a.cpu_seconds)
    kms, rays, trav = kernel_detail(w, a.integrator)
    if bdpt:
        metric = "path samples/s (BDPT, random shapes ~1M tris + 8 emissive triangles, 1280x720, depth 8)"
    else:
        metric = "path samples/s (NEE, random shapes ~1M tris, 1920x1080, depth 12)"
    detail = {"scene_build_s": w["t_scene"], "upload_and_bvh_s": w["t_upload"], "kernel_ms": kms, "rays": rays,
              "traversal": trav, "gpu_vs_cpu": (w["value"] / cpu["value"]) if cpu else None}
    out = {
        "metric": metric,
        "value": w["value"],
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": w["secs"] * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-shapes scene, seed 2)",
        "config": {"workload": "%s random-shapes %d shapes / %d tris, %dx%d, depth %d, 1 spp per step" %
                               (a.integrator.upper(), a.shapes, w["scene"].triangle_count, a.width, a.height, a.depth),
                   "launch": dict(w["cfg"], passes_in_flight=int(w["prof"][13]))},
        "roofline": roof,
        "cpu_baseline": cpu,
        "detail": detail,
    }
    del w
    if not a.no_extras:
        detail.update(extras(a))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
