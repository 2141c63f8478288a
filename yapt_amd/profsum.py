"""Summaries of rocprofv3 CSV output: per-kernel timing tables and HBM bytes per launch.

    python -m yapt_amd.profsum stats  <kernel-trace dir> "<command>" [rows]  > summary.md
    python -m yapt_amd.profsum traffic <fetch dir> <write dir> <calib fetch dir> <calib write dir> \
           <kernel substring> <key> <calib MiB> [traffic.json [summary.md "<command>"...]]
    python -m yapt_amd.profsum l2 <TCC_HIT/TCC_MISS dir> <kernel substring> ["<command>"...] > summary.md
    python -m yapt_amd.profsum sq <pmc dir> <kernel-trace dir> "<substr1;substr2;...>" ["<command>"...] > summary.md

``stats`` renders the ``*_kernel_stats.csv`` of a ``rocprofv3 --kernel-trace --stats`` run as a
markdown table. ``traffic`` turns the FETCH_SIZE / WRITE_SIZE counters (KiB per dispatch) of one
kernel into HBM bytes per launch, each scaled by what the same counter reads for a plain device copy
of known size (``yapt_amd.calib_copy``), and stores the result under ``key`` in traffic.json.
``l2`` gives the L2 (TCC) hit rate of a kernel from a ``--pmc TCC_HIT_sum TCC_MISS_sum`` run: the
share of its memory requests that never reach the fabric (MALL or HBM).
``sq`` tabulates, per kernel, the mean per dispatch of every counter of a ``--pmc`` run (rocprofv3's
derived VALUBusy, VALUUtilization, SALUBusy, OccupancyPercent next to raw SQ instruction counts) and the
register and LDS allocation that the kernel trace records.
``traffic`` takes ``--launch=<json>``: the launch configuration the counters were collected at.
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


def counters_by_kernel(dirname, kernel_substr):
    """mean value per dispatch of every counter recorded for kernels whose name contains kernel_substr"""
    acc = {}
    with open(_one(dirname, "counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if kernel_substr in r["Kernel_Name"]:
                acc.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: _mean(v) for k, v in acc.items()}, max((len(v) for v in acc.values()), default=0)


def resources_by_kernel(dirname, kernel_substr):
    """register and LDS allocation of the first dispatch of a kernel in a kernel-trace CSV"""
    with open(_one(dirname, "kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            if kernel_substr in r["Kernel_Name"]:
                return {k: r[k] for k in r if any(t in k for t in ("VGPR", "SGPR", "LDS", "Scratch", "Workgroup_Size_X"))}
    return {}


def sq_markdown(pmc_dir, trace_dir, substrs, commands):
    rows = []
    names = []
    for k in substrs:
        vals, n = counters_by_kernel(pmc_dir, k)
        res = resources_by_kernel(trace_dir, k) if trace_dir != "-" else {}
        rows.append((k, n, vals, res))
        for c in vals:
            if c not in names:
                names.append(c)
    out = ["# Issue and occupancy counters per kernel", "", "Commands:", "", "```"] + list(commands) + ["```", "",
           "Means per dispatch. VALUBusy, SALUBusy: % of GPU time the vector / scalar ALUs issue; VALUUtilization:"
           " % of active lanes in issued vector instructions; OccupancyPercent: resident waves as % of the maximum.",
           ""]
    out.append("| kernel | dispatches | " + " | ".join(names) + " | SQ_INSTS_SALU / SQ_INSTS_VALU | allocation |")
    out.append("|---" * (len(names) + 4) + "|")
    for k, n, vals, res in rows:
        cells = ["%.4g" % vals[c] if c in vals else "-" for c in names]
        ratio = "%.2f" % (vals["SQ_INSTS_SALU"] / vals["SQ_INSTS_VALU"]) if vals.get("SQ_INSTS_VALU") else "-"
        alloc = ", ".join("%s %s" % (a, b) for a, b in res.items()) or "-"
        out.append("| `%s` | %d | %s | %s | %s |" % (k, n, " | ".join(cells), ratio, alloc))
    return "\n".join(out) + "\n"


def main(argv):
    launch = None
    for x in list(argv):
        if x.startswith("--launch="):
            launch = json.loads(x.split("=", 1)[1])
            argv.remove(x)
    if argv[0] == "stats":
        sys.stdout.write(stats_markdown(argv[1], argv[2], int(argv[3]) if len(argv) > 3 else 12))
    elif argv[0] == "traffic":
        fetch_dir, write_dir, cfd, cwd, kern, key, mib = argv[1:8]
        path = argv[8] if len(argv) > 8 else os.path.join(os.path.dirname(os.path.dirname(
            os.path.abspath(__file__))), "profiles", "traffic.json")
        t = traffic(fetch_dir, write_dir, cfd, cwd, kern, float(mib))
        if launch is not None:
            t["launch"] = launch
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
    elif argv[0] == "sq":
        sys.stdout.write(sq_markdown(argv[1], argv[2], argv[3].split(";"), argv[4:]))
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
