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

DEFAULT = {"bvh_width": 2, "block_size": 128, "lds_stack_depth": 24}


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


def autotune(integrator, scene, frame, key, passes=4, widths=(2, 8), blocks=(64, 128, 256), lds_depths=(16, 24, 32),
             verbose=True):
    """Times `passes` full passes for every configuration and stores the fastest under `key`."""
    results = []
    for w in widths:
        for b in blocks:
            for d in lds_depths:
                integrator.set_option("bvh_width", w)
                integrator.set_option("block_size", b)
                integrator.set_option("lds_stack_depth", d)
                integrator.render(0, 1)
                integrator.synchronize()
                t = time.perf_counter()
                integrator.render(1, passes)
                integrator.synchronize()
                dt = (time.perf_counter() - t) / passes
                results.append({"config": {"bvh_width": w, "block_size": b, "lds_stack_depth": d}, "ms_per_pass": dt * 1e3})
                if verbose:
                    print("autotune %s w=%d block=%d lds=%d: %.2f ms/pass" % (key, w, b, d, dt * 1e3), flush=True)
    best = min(results, key=lambda r: r["ms_per_pass"])
    store = _load()
    store[key] = {"config": best["config"], "ms_per_pass": best["ms_per_pass"], "all": results}
    with open(TUNED, "w") as f:
        json.dump(store, f, indent=1, sort_keys=True)
    for k, v in best["config"].items():
        integrator.set_option(k, v)
    return best
