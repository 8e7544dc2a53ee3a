"""Summarise rocprofv3 output for one kernel into a short text report.

    python -m diaphane_amd.rocprof_summary --kernel k_fused_aligned \
        --trace DIR/kt_kernel_stats.csv --fetch DIR/pf_counter_collection.csv \
        --write DIR/pw_counter_collection.csv --cells N --bytes-per-cell 49 \
        --command "..." --out profiles/NAME.md [--json profiles/NAME.json]

FETCH_SIZE on gfx950 counts the TCC_EA0 read interface only; the factor that
turns it into bytes is measured on the same run from a kernel of known
traffic (``--calibrate-kernel`` with ``--calibrate-bytes``), the same way for
WRITE_SIZE. With no calibration kernel the factors are 1.
"""
from __future__ import annotations

import argparse
import csv
import json
import statistics
from typing import Dict, List, Optional


def _rows(path: str) -> List[Dict[str, str]]:
    with open(path, newline="") as handle:
        return list(csv.DictReader(handle))


def _match(name: str, needle: str) -> bool:
    return needle in name


def counter_values(path: str, kernel: str, counter: str) -> List[float]:
    return [float(r["Counter_Value"]) for r in _rows(path)
            if _match(r["Kernel_Name"], kernel) and r["Counter_Name"] == counter]


def display_name(traced: str) -> str:
    """A traced kernel name without its return type, namespaces and argument list."""
    name = traced.strip()
    if name.startswith("void "):
        name = name[len("void "):]
    depth = 0
    for i, ch in enumerate(name):  # cut at the argument list's opening parenthesis
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0 and i > 0 and not name[:i].endswith("::") and name[i - 1] != "(":
            if not name.startswith("anonymous", i + 1):
                name = name[:i]
                break
    return name.replace("dpa::(anonymous namespace)::", "").replace("dpa::", "")


def main(argv: Optional[List[str]] = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--trace", required=True, help="kernel_stats.csv from --kernel-trace --stats")
    ap.add_argument("--fetch", help="counter_collection.csv of a --pmc FETCH_SIZE run")
    ap.add_argument("--write", help="counter_collection.csv of a --pmc WRITE_SIZE run")
    ap.add_argument("--calibrate-kernel", help="kernel of known traffic in the same runs")
    ap.add_argument("--calibrate-read-bytes", type=float, default=0.0)
    ap.add_argument("--calibrate-write-bytes", type=float, default=0.0)
    ap.add_argument("--read-scale", type=float, default=0.0,
                    help="FETCH_SIZE factor measured on another run (used without a calibration kernel)")
    ap.add_argument("--write-scale", type=float, default=0.0)
    ap.add_argument("--cells", type=float, required=True)
    ap.add_argument("--bytes-per-cell", type=float, required=True, help="algorithmic bytes per cell-update")
    ap.add_argument("--steps-per-launch", type=int, default=1, help="cell-updates per cell in one launch")
    ap.add_argument("--peak-gbs", type=float, default=8000.0)
    ap.add_argument("--command", default="")
    ap.add_argument("--out", required=True)
    ap.add_argument("--json")
    ap.add_argument("--note", action="append", default=[], help="a line appended to the summary")
    args = ap.parse_args(argv)

    stats = [r for r in _rows(args.trace) if _match(r["Name"], args.kernel)]
    if not stats:
        raise SystemExit(f"{args.kernel} not in {args.trace}")
    calls = sum(int(r["Calls"]) for r in stats)
    total_ns = sum(float(r["TotalDurationNs"]) for r in stats)
    avg_ms = total_ns / calls / 1e6
    algorithmic = args.cells * args.bytes_per_cell * args.steps_per_launch
    achieved = algorithmic / (avg_ms / 1e3) / 1e9

    kernel_name = display_name(stats[0]["Name"]) if len({r["Name"] for r in stats}) == 1 else args.kernel
    summary = {
        "kernel": kernel_name,
        "launches": calls,
        "avg_ms": avg_ms,
        "algorithmic_bytes_per_launch": algorithmic,
        "steps_per_launch": args.steps_per_launch,
        "achieved_gbs": achieved,
        "peak_gbs": args.peak_gbs,
        "frac": achieved / args.peak_gbs,
    }
    lines = [f"# rocprofv3 summary: {kernel_name}", "", f"Command: `{args.command}`", "",
             "| quantity | value |", "|---|---|",
             f"| launches profiled | {calls} |",
             f"| average duration | {avg_ms:.3f} ms |",
             f"| algorithmic bytes / launch | {algorithmic / 1e9:.3f} GB ({args.bytes_per_cell:g} B × "
             + (f"{args.steps_per_launch} steps × " if args.steps_per_launch != 1 else "")
             + f"{args.cells:.0f} cells) |",
             f"| achieved (algorithmic / duration) | {achieved:.1f} GB/s = {100 * achieved / args.peak_gbs:.1f}% of {args.peak_gbs:.0f} GB/s |"]

    read_scale = args.read_scale or 1.0
    write_scale = args.write_scale or 1.0
    if args.calibrate_kernel and args.fetch and args.calibrate_read_bytes:
        raw = counter_values(args.fetch, args.calibrate_kernel, "FETCH_SIZE")
        if raw:
            read_scale = args.calibrate_read_bytes / (statistics.median(raw) * 1024.0)
    if args.calibrate_kernel and args.write and args.calibrate_write_bytes:
        raw = counter_values(args.write, args.calibrate_kernel, "WRITE_SIZE")
        if raw:
            write_scale = args.calibrate_write_bytes / (statistics.median(raw) * 1024.0)
    traffic = 0.0
    for label, path, counter, scale in (("read", args.fetch, "FETCH_SIZE", read_scale),
                                        ("write", args.write, "WRITE_SIZE", write_scale)):
        if not path:
            continue
        values = counter_values(path, args.kernel, counter)
        if not values:
            continue
        per_launch = statistics.median(values) * 1024.0 * scale
        traffic += per_launch
        summary[f"{label}_bytes_per_launch"] = per_launch
        summary[f"{label}_scale"] = scale
        lines.append(f"| L2→memory {label} / launch ({counter} × {scale:.3f}) | {per_launch / 1e9:.3f} GB |")
    if traffic:
        summary["traffic_bytes_per_launch"] = traffic
        summary["traffic_over_algorithmic"] = traffic / algorithmic
        lines.append(f"| L2→memory traffic / algorithmic | {traffic / algorithmic:.3f} |")
        lines.append(f"| L2→memory traffic rate | {traffic / (avg_ms / 1e3) / 1e9:.1f} GB/s |")
    if args.read_scale or args.write_scale:
        lines += ["", f"Counter scale factors ({read_scale:.3f} read, {write_scale:.3f} write) were measured "
                  f"with the stream-probe kernel of known traffic on the 768³ run."]
    if args.calibrate_kernel:
        lines += ["", f"Counter calibration: `{args.calibrate_kernel}` moves "
                  f"{args.calibrate_read_bytes / 1e9:.3f} GB read and {args.calibrate_write_bytes / 1e9:.3f} GB "
                  f"written per launch; the counters' ratio to that gives the scale factors above."]
    if args.note:
        lines += [""] + args.note
    with open(args.out, "w") as handle:
        handle.write("\n".join(lines) + "\n")
    if args.json:
        with open(args.json, "w") as handle:
            json.dump(summary, handle, indent=2)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
