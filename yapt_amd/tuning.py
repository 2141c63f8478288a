"""Launch-configuration autotuner and the store of tuned choices.

The autotuner searches BVH width (2 vs 8), threads per block, LDS traversal-stack depth, passes in
flight (queue depth), persistent vs one-workgroup-per-ray launches with their waves per CU and refill
threshold, the queue length below which rays are traced wave-per-ray instead of thread-per-ray, and
the shading launch size for a workload, timing whole passes, and records the fastest
configuration in ``tuned.json`` next to this file. ``best_config`` returns the stored choice (or
defaults).
"""
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNED = os.path.join(HERE, "tuned.json")
ROOT = os.path.dirname(HERE)

DEFAULT = {"bvh_width": 2, "block_size": 64, "lds_stack_depth": 16, "passes_in_flight": 0, "persistent": 1,
           "persistent_waves": 32, "refill_threshold": 16, "leaf_share": 2, "shade_waves": 64, "shade_reg_waves": 4,
           "wave_ray_below": 0, "steps_per_fetch": 8}


def _load(path=None):
    try:
        with open(path or TUNED) as f:
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
    """The record of a kernel's rocprofv3 FETCH_SIZE/WRITE_SIZE runs (profiles/traffic.json): HBM bytes per
    launch ("bytes_per_launch") and the launch shape it was profiled at ("launch"), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


SPACE = {
    "bvh_width": (2, 8),
    "block_size": (64, 128, 256),
    "lds_stack_depth": (8, 12, 16, 24),
    "passes_in_flight": (8, 16, 24, 32),
    "persistent": (0, 1),
    "persistent_waves": (8, 16, 24, 32),
    "refill_threshold": (4, 8, 16, 32, 64),
    "leaf_share": (1, 2, 3, 4),
    "shade_waves": (0, 16, 32, 64, 128),
    "shade_reg_waves": (2, 4, 8),
    "wave_ray_below": (0, 2048, 8192, 32768),
    "steps_per_fetch": (1, 4, 8, 16),
}

SPACE_BDPT = {
    "bvh_width": (2, 8),
    "bdpt_wavefront": (1, 0),
    "bdpt_waves": (1, 2, 3, 4, 6),
    "bdpt_split_connections": (0, 1, 2),
    "block_size": (64, 128, 256),
    "persistent_waves": (16, 24, 32),
    "refill_threshold": (8, 16, 32),
    "leaf_share": (1, 2, 4),
    "lds_stack_depth": (12, 16, 24),
    "steps_per_fetch": (1, 4, 8, 16),
}


def _time(integrator, cfg, passes):
    for k, v in cfg.items():
        integrator.set_option(k, v)
    integrator.render(0, 2)
    integrator.synchronize()
    t = time.perf_counter()
    integrator.render(2, passes)
    integrator.synchronize()
    return (time.perf_counter() - t) / passes * 1e3


def autotune(integrator, key, passes=16, space=None, rounds=2, start=None, verbose=True, out=None, history=None):
    """Coordinate descent over the launch options: starting from `start` (the stored or default
    configuration), each option in turn is set to every candidate of `space` with the others fixed,
    timing `passes` full passes (scene and frame already set), and the fastest value is kept;
    `rounds` sweeps over all options. Stores the best configuration under `key` and applies it; the
    timing of every configuration tried goes to the JSON file `history` when one is given."""
    space = space or SPACE
    cur = {k: v for k, v in dict(start or best_config(key)).items() if k in space or k in DEFAULT}
    for k in space:
        cur.setdefault(k, space[k][0])
    results = []
    seen = {}

    def measure(cfg):
        sig = tuple(sorted(cfg.items()))
        if sig not in seen:
            ms = _time(integrator, cfg, passes)
            seen[sig] = ms
            results.append({"config": dict(cfg), "ms_per_pass": ms})
            if verbose:
                print("autotune %s %s: %.3f ms/pass" % (key, cfg, ms), flush=True)
        return seen[sig]

    best_ms = measure(cur)
    for _ in range(rounds):
        changed = False
        for name, cands in space.items():
            for v in cands:
                if v == cur[name]:
                    continue
                cfg = dict(cur)
                cfg[name] = v
                ms = measure(cfg)
                if ms < best_ms * 0.995:
                    best_ms, cur, changed = ms, cfg, True
        if not changed:
            break
    store = _load(out)
    store[key] = {"config": cur, "ms_per_pass": best_ms}
    with open(out or TUNED, "w") as f:
        json.dump(store, f, indent=1, sort_keys=True)
    if history:
        with open(history, "w") as f:
            json.dump({key: results}, f, indent=1, sort_keys=True)
    for k, v in cur.items():
        integrator.set_option(k, v)
    return {"config": cur, "ms_per_pass": best_ms}
