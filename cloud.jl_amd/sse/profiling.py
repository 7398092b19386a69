"""Summaries of rocprofv3 output for the residual kernels.

Two summaries, both plain text or JSON so they can be kept beside the code:

* ``kernel-stats``: per-kernel dispatch count and mean/min/max duration from a
  ``rocprofv3 --kernel-trace --stats --output-format csv`` directory
  (``*_kernel_stats.csv``, falling back to ``*_kernel_trace.csv``);
* ``traffic``: HBM bytes per dispatch from two ``rocprofv3 --pmc`` runs, one
  collecting ``FETCH_SIZE`` and one ``WRITE_SIZE`` (KiB per dispatch), each
  scaled by what the same counter reads for an elementwise copy kernel of known
  size (``bench.py --copy-probe-mib``), so that the figure is in real bytes
  whatever fraction of the memory channels the counter samples.

* ``counters``: every counter of one or more ``rocprofv3 --pmc`` runs, mean per
  dispatch, for each kernel slot.

Usage::

    python -m sse.profiling kernel-stats <trace_dir>
    python -m sse.profiling counters <pmc_dir> [<pmc_dir> ...]
    python -m sse.profiling traffic <fetch_dir> <write_dir> --copy-mib 256 \
        --workload NAME [--json profiles/hbm_traffic.json]
"""

from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import re
import statistics
from collections import defaultdict

# kernel slot name -> regex on the demangled kernel name
SLOTS = {
    "nodal_values": re.compile(r"k_nodal_values"),
    "time_derivative": re.compile(r"k_time_derivative"),
    "copy_probe": re.compile(r"elementwise_kernel"),
}


def short_name(kernel_name: str) -> str:
    name = kernel_name.replace("(anonymous namespace)::", "").split("(")[0]
    return name.replace("void ", "").replace("sse::", "")


def slot_of(kernel_name: str):
    for slot, rx in SLOTS.items():
        if rx.search(kernel_name):
            return slot
    return None


def _csv_rows(directory: str, suffix: str):
    files = sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True))
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def kernel_stats(directory: str):
    """Per-kernel {calls, mean_us, min_us, max_us} from a kernel-trace run."""
    trace = _csv_rows(directory, "kernel_trace.csv")
    out = {}
    if trace:
        durs = defaultdict(list)
        for r in trace:
            durs[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
        for k, v in durs.items():
            out[k] = {"calls": len(v), "mean_us": statistics.fmean(v), "min_us": min(v),
                      "max_us": max(v), "total_us": sum(v)}
        return out
    for r in _csv_rows(directory, "kernel_stats.csv"):
        out[r["Name"]] = {"calls": int(r["Calls"]), "mean_us": float(r["AverageNs"]) * 1e-3,
                          "min_us": float(r["MinNs"]) * 1e-3, "max_us": float(r["MaxNs"]) * 1e-3,
                          "total_us": float(r["TotalDurationNs"]) * 1e-3}
    return out


def format_kernel_stats(stats) -> str:
    lines = [f"{'kernel':60s} {'calls':>6s} {'mean us':>10s} {'min us':>10s} {'max us':>10s}"]
    for k, s in sorted(stats.items(), key=lambda kv: -kv[1]["total_us"]):
        lines.append(f"{short_name(k)[:60]:60s} {s['calls']:6d} {s['mean_us']:10.2f} "
                     f"{s['min_us']:10.2f} {s['max_us']:10.2f}")
    return "\n".join(lines)


def counter_means(directory: str, counter: str):
    """Mean counter value per dispatch (summed over the rows of one dispatch) by kernel name."""
    per_dispatch = defaultdict(float)
    names = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        key = r["Dispatch_Id"]
        per_dispatch[key] += float(r["Counter_Value"])
        names[key] = r["Kernel_Name"]
    by_kernel = defaultdict(list)
    for key, v in per_dispatch.items():
        by_kernel[names[key]].append(v)
    return {k: statistics.fmean(v) for k, v in by_kernel.items()}


def all_counters(directories):
    """{slot: {counter: mean per dispatch}} over the counter-collection runs in `directories`."""
    per = defaultdict(lambda: defaultdict(float))  # (dir, dispatch) -> counter -> value
    names = {}
    for d in directories:
        for r in _csv_rows(d, "counter_collection.csv"):
            key = (d, r["Dispatch_Id"])
            per[key][r["Counter_Name"]] += float(r["Counter_Value"])
            names[key] = r["Kernel_Name"]
    acc = defaultdict(lambda: defaultdict(list))
    for key, cs in per.items():
        slot = slot_of(names[key])
        if slot is None:
            continue
        for c, v in cs.items():
            acc[slot][c].append(v)
    return {s: {c: statistics.fmean(v) for c, v in cs.items()} for s, cs in acc.items()}


def format_counters(data) -> str:
    slots = [s for s in ("nodal_values", "time_derivative") if s in data]
    names = sorted({c for s in slots for c in data[s]})
    lines = [f"{'counter (mean per dispatch)':34s}" + "".join(f"{s:>18s}" for s in slots)]
    for c in names:
        lines.append(f"{c:34s}" + "".join(f"{data[s].get(c, float('nan')):18.6g}" for s in slots))
    return "\n".join(lines)


def traffic(fetch_dir: str, write_dir: str, copy_mib: int):
    """Calibrated HBM bytes per dispatch for each kernel slot."""
    fetch = counter_means(fetch_dir, "FETCH_SIZE")
    write = counter_means(write_dir, "WRITE_SIZE")
    copy_bytes = copy_mib * (1 << 20)

    def by_slot(d):
        out = {}
        for k, v in d.items():
            s = slot_of(k)
            if s is not None:
                out.setdefault(s, []).append(v)
        return {s: max(v) for s, v in out.items()}  # the copy probe is the largest elementwise kernel

    f, w = by_slot(fetch), by_slot(write)
    if "copy_probe" not in f or "copy_probe" not in w:
        raise RuntimeError("copy probe kernel not found in the counter runs")
    f_scale = copy_bytes / (f["copy_probe"] * 1024.0)
    w_scale = copy_bytes / (w["copy_probe"] * 1024.0)
    out = {"calibration": {"copy_bytes": copy_bytes,
                           "fetch_kib_read_for_copy": f["copy_probe"],
                           "write_kib_read_for_copy": w["copy_probe"],
                           "fetch_scale": f_scale, "write_scale": w_scale}}
    for s in f:
        if s == "copy_probe" or s not in w:
            continue
        rb = f[s] * 1024.0 * f_scale
        wb = w[s] * 1024.0 * w_scale
        out[s] = {"fetch_kib_raw": f[s], "write_kib_raw": w[s], "read_bytes": rb,
                  "write_bytes": wb, "bytes": rb + wb}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m sse.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("kernel-stats")
    a.add_argument("dir")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--copy-mib", type=int, required=True)
    b.add_argument("--workload", required=True)
    b.add_argument("--json", default=None, help="merge the result into this JSON file")
    args = ap.parse_args(argv)
    if args.cmd == "kernel-stats":
        print(format_kernel_stats(kernel_stats(args.dir)))
        return
    if args.cmd == "counters":
        print(format_counters(all_counters(args.dirs)))
        return
    t = traffic(args.fetch_dir, args.write_dir, args.copy_mib)
    print(json.dumps(t, indent=1))
    if args.json:
        data = {}
        if os.path.exists(args.json):
            with open(args.json) as fh:
                data = json.load(fh)
        data[args.workload] = {k: (v["bytes"] if k != "calibration" else v) for k, v in t.items()}
        with open(args.json, "w") as fh:
            json.dump(data, fh, indent=1)


if __name__ == "__main__":
    main()
