"""Summaries of rocprofv3 CSV output: per-kernel timing tables and HBM bytes per launch.

    python -m yapt_amd.profsum stats  <kernel-trace dir> "<command>" [rows]  > summary.md
    python -m yapt_amd.profsum traffic <fetch dir> <write dir> <calib fetch dir> <calib write dir> \
           <kernel substring> <key> <calib MiB> [traffic.json [summary.md "<command>"...]]
    python -m yapt_amd.profsum l2 <TCC_HIT/TCC_MISS dir> <kernel substring> ["<command>"...] > summary.md

``stats`` renders the ``*_kernel_stats.csv`` of a ``rocprofv3 --kernel-trace --stats`` run as a
markdown table. ``traffic`` turns the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) of one
kernel into HBM bytes per launch, each scaled by what the same counter reads for a plain device copy
of known size (``yapt_amd.calib_copy``), and stores the result under ``key`` in traffic.json.
``l2`` gives the L2 (TCC) hit rate of a kernel from a ``--pmc TCC_HIT_sum TCC_MISS_sum`` run: the
share of its memory requests that never reach the fabric (MALL or HBM).
"""
import csv
import glob
import json
import os
import sys


def _one(dirname, suffix):
    files = sorted(glob.glob(os.path.join(dirname, "**", "*" + suffix), recursive=True))
    if not files:
        raise FileNotFoundError("no *%s under %s" % (suffix, dirname))
    return files[0]


def kernel_stats(dirname):
    rows = []
    with open(_one(dirname, "kernel_stats.csv")) as f:
        for r in csv.DictReader(f):
            rows.append({"name": r["Name"], "calls": int(r["Calls"]), "total_ns": float(r["TotalDurationNs"]),
                         "avg_ns": float(r["AverageNs"]), "pct": float(r["Percentage"]),
                         "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"])})
    return rows


def stats_markdown(dirname, command, top=12):
    rows = kernel_stats(dirname)
    out = ["# rocprofv3 kernel statistics", "", "Command:", "", "```", command, "```", "",
           "| kernel | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        name = r["name"] if len(r["name"]) < 90 else r["name"][:87] + "..."
        out.append("| `%s` | %d | %.2f | %.1f | %.1f | %.1f | %.2f |" % (
            name.replace("|", "\\|"), r["calls"], r["total_ns"] * 1e-6, r["avg_ns"] * 1e-3,
            r["min_ns"] * 1e-3, r["max_ns"] * 1e-3, r["pct"]))
    return "\n".join(out) + "\n"


def counter_values(dirname, counter, kernel_substr):
    vals = []
    with open(_one(dirname, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter and kernel_substr in r["Kernel_Name"]:
                vals.append(float(r["Counter_Value"]))
    return vals


def _mean(v):
    if not v:
        raise ValueError("no matching dispatches")
    return sum(v) / len(v)


def traffic(fetch_dir, write_dir, cfetch_dir, cwrite_dir, kernel_substr, calib_mib):
    """Mean HBM bytes per launch of the kernel: counters scaled by the copy calibration."""
    copy_bytes = calib_mib * float(1 << 20)
    cf = max(counter_values(cfetch_dir, "FETCH_SIZE", "copyBuffer"))
    cw = max(counter_values(cwrite_dir, "WRITE_SIZE", "copyBuffer"))
    fetch_factor = copy_bytes / (cf * 1024.0)     # true bytes per counted byte
    write_factor = copy_bytes / (cw * 1024.0)
    fv = counter_values(fetch_dir, "FETCH_SIZE", kernel_substr)
    wv = counter_values(write_dir, "WRITE_SIZE", kernel_substr)
    fetch = _mean(fv) * 1024.0 * fetch_factor
    write = _mean(wv) * 1024.0 * write_factor
    return {"bytes_per_launch": fetch + write, "fetch_bytes": fetch, "write_bytes": write,
            "fetch_factor": fetch_factor, "write_factor": write_factor, "dispatches": [len(fv), len(wv)],
            "kernel": kernel_substr}


def traffic_markdown(t, key, commands):
    out = ["# HBM traffic per launch: %s (`%s`)" % (key, t["kernel"]), "", "Commands:", "", "```"] + list(commands) + [
        "```", "",
        "| quantity | value |", "|---|---|",
        "| dispatches averaged (FETCH, WRITE) | %d, %d |" % tuple(t["dispatches"]),
        "| FETCH_SIZE scale (copy bytes / counted bytes) | %.4f |" % t["fetch_factor"],
        "| WRITE_SIZE scale (copy bytes / counted bytes) | %.4f |" % t["write_factor"],
        "| fetched bytes per launch | %.4g |" % t["fetch_bytes"],
        "| written bytes per launch | %.4g |" % t["write_bytes"],
        "| HBM bytes per launch | %.4g |" % t["bytes_per_launch"]]
    return "\n".join(out) + "\n"


def l2_hit_rate(dirname, kernel_substr):
    hits = counter_values(dirname, "TCC_HIT_sum", kernel_substr)
    miss = counter_values(dirname, "TCC_MISS_sum", kernel_substr)
    h, m = sum(hits), sum(miss)
    return {"hit_rate": h / (h + m) if h + m else 0.0, "hits_per_launch": _mean(hits),
            "misses_per_launch": _mean(miss), "dispatches": [len(hits), len(miss)], "kernel": kernel_substr}


def l2_markdown(r, commands):
    out = ["# L2 hit rate: `%s`" % r["kernel"], "", "Commands:", "", "```"] + list(commands) + ["```", "",
           "| quantity | value |", "|---|---|",
           "| dispatches averaged (HIT, MISS) | %d, %d |" % tuple(r["dispatches"]),
           "| L2 hits per launch | %.4g |" % r["hits_per_launch"],
           "| L2 misses per launch | %.4g |" % r["misses_per_launch"],
           "| L2 hit rate | %.3f |" % r["hit_rate"]]
    return "\n".join(out) + "\n"


def main(argv):
    if argv[0] == "stats":
        sys.stdout.write(stats_markdown(argv[1], argv[2], int(argv[3]) if len(argv) > 3 else 12))
    elif argv[0] == "traffic":
        fetch_dir, write_dir, cfd, cwd, kern, key, mib = argv[1:8]
        path = argv[8] if len(argv) > 8 else os.path.join(os.path.dirname(os.path.dirname(
            os.path.abspath(__file__))), "profiles", "traffic.json")
        t = traffic(fetch_dir, write_dir, cfd, cwd, kern, float(mib))
        try:
            with open(path) as f:
                store = json.load(f)
        except (OSError, ValueError):
            store = {}
        store[key] = t
        with open(path, "w") as f:
            json.dump(store, f, indent=1, sort_keys=True)
        print(json.dumps(t))
        if len(argv) > 9:
            with open(argv[9], "w") as f:
                f.write(traffic_markdown(t, key, argv[10:]))
    elif argv[0] == "l2":
        r = l2_hit_rate(argv[1], argv[2])
        sys.stdout.write(l2_markdown(r, argv[3:]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
