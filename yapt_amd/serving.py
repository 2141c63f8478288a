"""Serving workload: a seeded Poisson stream of render jobs over resident scenes, submitted to the
render server at their arrival times; reports p50/p99 latency (submit to film complete) and
throughput.

    python -m yapt_amd.serving [--rate 40] [--jobs 120] [--budget-gb 96] [--max-active 8] [--no-fuse]

Default mix: 4 resident random-shapes scenes (~10k, ~100k, ~1M, ~5M triangles), resolution uniform in
{256x256, 512x512, 1280x720, 1920x1080}, spp uniform in {1, 4, 16, 64}, integrator BPT : NEE : BDPT
= 1 : 2 : 1.
"""
import argparse
import json
import sys
import threading
import time

import numpy as np

RESOLUTIONS = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
SPP = [1, 4, 16, 64]
INTEGRATORS = [0, 1, 1, 2]            # BPT : NEE : BDPT = 1 : 2 : 1
SCENE_SHAPES = [20, 200, 2000, 10000]  # ~10k, ~100k, ~1M, ~5M triangles


def make_requests(n, rate, seed=4, n_scenes=4):
    """n requests: Poisson arrival times (exponential gaps, mean 1/rate) and a uniform draw of
    scene, resolution, spp and integrator (the integrator list carries the 1:2:1 weights)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        out.append({"t": float(t[i]), "scene": int(rng.integers(n_scenes)),
                    "res": RESOLUTIONS[int(rng.integers(len(RESOLUTIONS)))],
                    "spp": SPP[int(rng.integers(len(SPP)))],
                    "integrator": INTEGRATORS[int(rng.integers(len(INTEGRATORS)))]})
    return out


def percentiles(latencies, qs=(50, 90, 99)):
    a = np.sort(np.asarray(latencies, np.float64))
    return {"p%d" % q: float(np.percentile(a, q)) if len(a) else None for q in qs}


def run(server, scenes, cams, requests, depth=8, keep_films=()):
    """Submits each request at its arrival time (relative to now) from a feeder thread; returns one
    record per request with its latency (and its film, for the request indices in `keep_films`)."""
    from . import camera
    t0 = time.perf_counter()
    ids = [None] * len(requests)

    def feeder():
        for i, r in enumerate(requests):
            d = r["t"] - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            w, h = r["res"]
            fr = camera.make_frame(cams[r["scene"]], w, h, max_depth=depth)
            ids[i] = server.submit(scenes[r["scene"]], r["integrator"], fr, r["spp"])

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    recs = []
    for i, r in enumerate(requests):
        server.wait(ids[i])
        film, times = server.take(ids[i])
        recs.append(dict(r, latency=times[2] - times[0], queue=times[1] - times[0],
                         samples=r["res"][0] * r["res"][1] * r["spp"], finish=times[2], submit=times[0],
                         ok=bool(np.isfinite(film).all()), film=film if i in keep_films else None))
    return recs


def trace(jobs=120, rate=40.0, shapes=SCENE_SHAPES, budget_gb=96.0, max_active=8, seed=4, log=lambda m: None, fuse=True):
    """Builds the resident scenes, replays a seeded Poisson stream of `jobs` requests at `rate` per second
    through a render server with the tuned launch options, and returns the latency and throughput
    figures (seconds from submit to film complete)."""
    from . import _lib, scenes as S, tuning
    res_scenes, cams, tris = [], [], []
    t_build = time.perf_counter()
    for k, n in enumerate(shapes):
        sc, cam = S.random_shapes(n_shapes=n, seed=2 + k, emissive_triangles=8)
        res_scenes.append(_lib.ResidentScene(sc))
        cams.append(cam)
        tris.append(sc.triangle_count)
        log("scene %d: %d triangles resident" % (k, sc.triangle_count))
    t_build = time.perf_counter() - t_build
    server = _lib.Server(0, int(budget_gb * (1 << 30)), max_active)
    server.set_option("fuse_jobs", 1 if fuse else 0)
    skip = ("passes_in_flight",)   # the server sizes batches per job
    for integ, key in ((0, "nee-shapes2000-1920x1080-d12"), (1, "nee-shapes2000-1920x1080-d12"),
                       (2, "bdpt-shapes2000-1280x720-d8")):
        for k, v in tuning.best_config(key).items():
            if k not in skip:
                server.set_option(k, v, integ)
    reqs = make_requests(jobs, rate, seed, len(shapes))
    t = time.perf_counter()
    recs = run(server, res_scenes, cams, reqs)
    wall = time.perf_counter() - t
    st = server.stats()
    fs = server.fused_stats()
    server.close()
    for r in res_scenes:
        r.close()
    lat = [r["latency"] for r in recs]
    first = min(r["submit"] for r in recs)
    last = max(r["finish"] for r in recs)
    return {
        "metric": "render-job latency (Poisson arrivals, continuous batching)",
        "rate_req_s": rate, "jobs": len(recs), "scene_triangles": tris, "scene_build_s": t_build,
        "latency_s": percentiles(lat), "mean_latency_s": float(np.mean(lat)),
        "queue_wait_s": percentiles([r["queue"] for r in recs]),
        "throughput_jobs_s": len(recs) / (last - first),
        "throughput_path_samples_s": sum(r["samples"] for r in recs) / (last - first),
        "offered_path_samples_s": sum(r["samples"] for r in recs) / max(1e-9, reqs[-1]["t"]),
        "wall_s": wall, "peak_active": st["peak_active"], "max_active": max_active,
        "fused_steps": fs["steps"], "jobs_per_fused_step": fs["job_chunks"] / fs["steps"] if fs["steps"] else None,
        "budget_gb": budget_gb, "all_finite": all(r["ok"] for r in recs),
        "by_integrator": {name: percentiles([r["latency"] for r in recs if r["integrator"] == k])
                          for k, name in ((0, "bpt"), (1, "nee"), (2, "bdpt"))},
    }


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--rate", type=float, default=40.0)
    p.add_argument("--jobs", type=int, default=120)
    p.add_argument("--budget-gb", type=float, default=96.0)
    p.add_argument("--max-active", type=int, default=8)
    p.add_argument("--shapes", default=",".join(str(s) for s in SCENE_SHAPES))
    p.add_argument("--seed", type=int, default=4)
    p.add_argument("--no-fuse", action="store_true", help="every job on its own stream (no fused steps)")
    a = p.parse_args()
    out = trace(a.jobs, a.rate, [int(x) for x in a.shapes.split(",")], a.budget_gb, a.max_active, a.seed,
                log=lambda m: print(m, file=sys.stderr, flush=True), fuse=not a.no_fuse)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
