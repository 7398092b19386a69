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

* ``busy``: how much of the window spanned by the large dispatches the GPU spends
  running kernels (the gaps between the dispatches of a graph replay);
* ``counters``: every counter of one or more ``rocprofv3 --pmc`` runs, mean per
  dispatch, for each kernel slot; with ``--elements`` and ``--trace``, the derived
  figures below the table (instructions per element, the VALU issue floor against
  the kernel's mean duration, the share of wave time spent waiting).

Usage::

    python -m sse.profiling kernel-stats <trace_dir>
    python -m sse.profiling counters <pmc_dir> [<pmc_dir> ...] [--elements N --trace <trace_dir>]
    python -m sse.profiling busy <trace_dir> --min-grid THREADS
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
    "vjp_v1": re.compile(r"k_vjp_v1"),
    "vjp_v2": re.compile(r"k_vjp_v2"),
    "vjp_v3": re.compile(r"k_vjp_v3"),
    "copy_probe": re.compile(r"elementwise_kernel"),
}
# the slots the counter tables show, in this order
TABLE_SLOTS = ("nodal_values", "time_derivative", "vjp_v1", "vjp_v2", "vjp_v3")


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


def kernel_stats(directory: str, last: int = 0):
    """Per-kernel {calls, mean_us, min_us, max_us} from a kernel-trace run; with `last`, over
    each kernel's last `last` dispatches only (in launch order)."""
    trace = _csv_rows(directory, "kernel_trace.csv")
    out = {}
    if trace:
        durs = defaultdict(list)
        for r in sorted(trace, key=lambda r: int(r["Start_Timestamp"])):
            durs[r["Kernel_Name"]].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
        if last > 0:
            durs = {k: v[-last:] for k, v in durs.items()}
        for k, v in durs.items():
            out[k] = {"calls": len(v), "mean_us": statistics.fmean(v), "min_us": min(v),
                      "max_us": max(v), "total_us": sum(v)}
        return out
    for r in _csv_rows(directory, "kernel_stats.csv"):
        out[r["Name"]] = {"calls": int(r["Calls"]), "mean_us": float(r["AverageNs"]) * 1e-3,
                          "min_us": float(r["MinNs"]) * 1e-3, "max_us": float(r["MaxNs"]) * 1e-3,
                          "total_us": float(r["TotalDurationNs"]) * 1e-3}
    return out


def kernel_stats_by_grid(directory: str):
    """Per (kernel, grid size in threads) {calls, mean_us, min_us, max_us} from a kernel-trace
    run: one line per problem size when a run times several."""
    out = defaultdict(list)
    for r in _csv_rows(directory, "kernel_trace.csv"):
        key = f"{short_name(r['Kernel_Name'])} [grid {int(r['Grid_Size_X'])}]"
        out[key].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    return {k: {"calls": len(v), "mean_us": statistics.fmean(v), "min_us": min(v), "max_us": max(v),
                "total_us": sum(v)} for k, v in out.items()}


def busy(directory: str, min_grid: int):
    """GPU occupancy of the window spanned by the dispatches whose grid has at least
    `min_grid` threads: the union of all kernels' busy intervals in that window against its
    length, the number of dispatches in it, and the gaps between consecutive dispatches."""
    rows = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), int(r["Grid_Size_X"]), r["Kernel_Name"])
            for r in _csv_rows(directory, "kernel_trace.csv")]
    big = [r for r in rows if r[2] >= min_grid]
    if not big:
        raise RuntimeError("no dispatch with that grid size")
    t0, t1 = min(r[0] for r in big), max(r[1] for r in big)
    win = sorted((a, b) for a, b, _, _ in rows if a >= t0 and b <= t1)
    busy_ns, cur_a, cur_b = 0, None, None
    gaps = []
    for a, b in win:
        if cur_b is None:
            cur_a, cur_b = a, b
        elif a > cur_b:
            busy_ns += cur_b - cur_a
            gaps.append(a - cur_b)
            cur_a, cur_b = a, b
        else:
            cur_b = max(cur_b, b)
    busy_ns += cur_b - cur_a
    span = t1 - t0
    return {"window_us": span * 1e-3, "busy_us": busy_ns * 1e-3, "busy_frac": busy_ns / span,
            "dispatches": len(win), "gaps": len(gaps), "mean_gap_us": (sum(gaps) / len(gaps) * 1e-3) if gaps else 0.0,
            "kernel_sum_us": sum(b - a for a, b in win) * 1e-3}


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
    slots = [s for s in TABLE_SLOTS if s in data]
    names = sorted({c for s in slots for c in data[s]})
    lines = [f"{'counter (mean per dispatch)':34s}" + "".join(f"{s:>18s}" for s in slots)]
    for c in names:
        lines.append(f"{c:34s}" + "".join(f"{data[s].get(c, float('nan')):18.6g}" for s in slots))
    return "\n".join(lines)


# MI355X: 256 CUs of 4 SIMDs, 2.4 GHz peak engine clock. A wave64 VALU instruction
# occupies its SIMD's 16 lanes for 4 cycles (FP64 at full rate); FP64 transcendental
# instructions (v_rcp_f64, v_sqrt_f64, ...) run at a quarter of that rate.
CUS, SIMDS_PER_CU, CLOCK_GHZ = 256, 4, 2.4


def derived(data, elements: int, mean_us=None):
    """{slot: {figure: value}} from the counter means of `data` (see ``all_counters``):
    instructions per element, the FP64 lane flops the VALU executed, the VALU issue
    floor (every SIMD issuing a VALU instruction every 4 cycles, 16 for FP64
    transcendentals) and its share of the mean duration, and the shares of wave time
    spent issuing VALU instructions and waiting."""
    out = {}
    for slot, c in data.items():
        if slot == "copy_probe":
            continue
        g = c.get
        d = {}
        if g("SQ_WAVES"):
            d["waves"] = g("SQ_WAVES")
        if g("SQ_INSTS_VALU") is not None:
            d["valu_insts_per_element"] = g("SQ_INSTS_VALU") / elements
            if g("SQ_WAVES"):
                d["valu_insts_per_wave"] = g("SQ_INSTS_VALU") / g("SQ_WAVES")
        for k in ("FMA_F64", "ADD_F64", "MUL_F64", "TRANS_F64", "INT32"):
            if g(f"SQ_INSTS_VALU_{k}") is not None:
                d[f"valu_{k.lower()}_per_element"] = g(f"SQ_INSTS_VALU_{k}") / elements
        if all(g(f"SQ_INSTS_VALU_{k}") is not None for k in ("FMA_F64", "ADD_F64", "MUL_F64")):
            d["fp64_lane_flops_per_element"] = 64 * (2 * g("SQ_INSTS_VALU_FMA_F64") + g("SQ_INSTS_VALU_ADD_F64")
                                                     + g("SQ_INSTS_VALU_MUL_F64")) / elements
        for k in ("SQ_INSTS_LDS", "SQ_INSTS_SALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR"):
            if g(k) is not None:
                d[k[3:].lower() + "_per_element"] = g(k) / elements
        if g("SQ_INSTS_VALU") is not None:
            cycles = 4 * (g("SQ_INSTS_VALU") + 3 * (g("SQ_INSTS_VALU_TRANS_F64") or 0.0))
            floor_us = cycles / (CUS * SIMDS_PER_CU) / (CLOCK_GHZ * 1e3)
            d["valu_issue_floor_us"] = floor_us
            if mean_us and slot in mean_us:
                d["mean_us"] = mean_us[slot]
                d["valu_issue_floor_frac"] = floor_us / mean_us[slot]
        if g("SQ_WAVE_CYCLES"):
            if g("SQ_WAIT_ANY") is not None:
                d["wave_time_waiting"] = g("SQ_WAIT_ANY") / g("SQ_WAVE_CYCLES")
            if g("SQ_ACTIVE_INST_VALU") is not None:
                d["wave_time_issuing_valu"] = g("SQ_ACTIVE_INST_VALU") / g("SQ_WAVE_CYCLES")
        if g("SQ_INSTS_LDS") and g("SQ_LDS_BANK_CONFLICT") is not None:
            d["lds_bank_conflict_cycles_per_lds_inst"] = g("SQ_LDS_BANK_CONFLICT") / g("SQ_INSTS_LDS")
        out[slot] = d
    return out


def slot_means(trace_dir: str):
    """{slot: mean duration in us} from a kernel-trace run (the slowest kernel of each slot)."""
    out = {}
    for k, st in kernel_stats(trace_dir).items():
        s = slot_of(k)
        if s is not None:
            out[s] = max(out.get(s, 0.0), st["mean_us"])
    return out


def format_derived(d) -> str:
    slots = [s for s in TABLE_SLOTS if s in d]
    names = sorted({k for s in slots for k in d[s]})
    lines = [f"{'derived':40s}" + "".join(f"{s:>18s}" for s in slots)]
    for k in names:
        lines.append(f"{k:40s}" + "".join(f"{d[s].get(k, float('nan')):18.6g}" for s in slots))
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
    a.add_argument("--by-grid", action="store_true", help="one line per (kernel, grid size)")
    a.add_argument("--last", type=int, default=0, help="each kernel's last N dispatches only")
    bz = sub.add_parser("busy")
    bz.add_argument("dir")
    bz.add_argument("--min-grid", type=int, required=True, help="threads; dispatches at least this large span the window")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--elements", type=int, default=0, help="elements per dispatch (derived figures)")
    c.add_argument("--trace", default=None, help="kernel-trace run of the same command (mean durations)")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_dir")
    b.add_argument("write_dir")
    b.add_argument("--copy-mib", type=int, required=True)
    b.add_argument("--workload", required=True)
    b.add_argument("--json", default=None, help="merge the result into this JSON file")
    args = ap.parse_args(argv)
    if args.cmd == "kernel-stats":
        print(format_kernel_stats(kernel_stats_by_grid(args.dir) if args.by_grid else kernel_stats(args.dir, args.last)))
        return
    if args.cmd == "busy":
        print(json.dumps(busy(args.dir, args.min_grid), indent=1))
        return
    if args.cmd == "counters":
        data = all_counters(args.dirs)
        print(format_counters(data))
        if args.elements:
            print()
            print(format_derived(derived(data, args.elements, slot_means(args.trace) if args.trace else None)))
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
