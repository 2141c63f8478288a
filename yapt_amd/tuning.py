"""Launch-configuration autotuner and the store of tuned choices.

The autotuner sweeps BVH width (2 vs 8), threads per block and LDS traversal-stack depth for a
workload, timing whole passes, and records the fastest configuration in ``tuned.json`` next to this
file. ``best_config`` returns the stored choice (or defaults).
"""
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned.json")
ROOT = os.path.dirname(HERE)

DEFAULT = {"bvh_width": 2, "block_size": 64, "lds_stack_depth": 16, "passes_in_flight": 0}


def _load():
    try:
        with open(TUNED) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def best_config(key):
    d = _load().get(key)
    cfg = dict(DEFAULT)
    if d:
        cfg.update(d.get("config", {}))
    return cfg


def measured_traffic(kernel_key):
    """HBM bytes per launch of a kernel from the rocprofv3 FETCH_SIZE/WRITE_SIZE runs (profiles/)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            v = json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None
    return None if v is None else v.get("bytes_per_launch")


def autotune(integrator, key, passes=16, widths=(2, 8), blocks=(64, 128), lds_depths=(8, 16, 24), flights=(8, 16),
             verbose=True, out=None):
    """Times `passes` full passes (scene and frame already set) for every configuration of BVH width,
    threads per block, LDS stack depth and passes in flight; stores the fastest under `key`."""
    results = []
    for g in flights:
        for w in widths:
            for b in blocks:
                for d in lds_depths:
                    cfg = {"bvh_width": w, "block_size": b, "lds_stack_depth": d, "passes_in_flight": g}
                    for k, v in cfg.items():
                        integrator.set_option(k, v)
                    integrator.render(0, 2)
                    integrator.synchronize()
                    t = time.perf_counter()
                    integrator.render(2, passes)
                    integrator.synchronize()
                    dt = (time.perf_counter() - t) / passes
                    results.append({"config": cfg, "ms_per_pass": dt * 1e3})
                    if verbose:
                        print("autotune %s %s: %.2f ms/pass" % (key, cfg, dt * 1e3), flush=True)
    best = min(results, key=lambda r: r["ms_per_pass"])
    store = _load()
    store[key] = {"config": best["config"], "ms_per_pass": best["ms_per_pass"], "all": results}
    with open(out or TUNED, "w") as f:
        json.dump(store, f, indent=1, sort_keys=True)
    for k, v in best["config"].items():
        integrator.set_option(k, v)
    return best
