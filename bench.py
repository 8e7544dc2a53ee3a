#!/usr/bin/env python3
"""Benchmark: Yee cell-updates/s of the fused gfx950 step, against the HBM
roofline, with the CPU oracle timed beside it.

Default workload (``layered_768``): a 768³ uniform grid (452,984,832 cells),
40 random dielectric z-slabs indexed through u8 into 16 materials, a centred
z-dipole, 1000 steps with |E|² accumulated (f64) over the last 500. Fields and
materials are resident in HBM before the timed region starts.

One step = one full leapfrog step of the whole grid (H, injections, E,
injections, intensity). Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
BYTES_PER_CELL = 49            # 6 f32 fields read+written + 1 u8 material index
BYTES_PER_CELL_INTENSITY = 65  # + f64 intensity read-modify-write


def cpu_baseline(workload_name: str, seconds: float = 10.0) -> dict:
    """Times the C oracle (a restatement of the reference's cpu.rs) on the
    host's cores over a bounded sample of the same workload."""
    import numpy as np  # noqa: F401

    import oracle
    from diaphane_amd import workloads

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    if workload_name == "layered_768":
        sample = workloads.layered_768(nz=64)
        label = "768x768x64 slab of the layered_768 stack"
    else:
        sample = workloads.WORKLOADS[workload_name]()
        label = f"{workload_name} (full grid)"
    sim = oracle.Simulation(sample.scene, threads=threads)
    sim.advance_by(2)  # warm-up
    steps = 0
    start = time.perf_counter()
    while True:
        sim.advance_by(1)
        steps += 1
        elapsed = time.perf_counter() - start
        if elapsed >= seconds and steps >= 3:
            break
    rate = sim.cells * steps / elapsed
    return {
        "value": rate,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{label}: {steps} steps after 2 warm-up, {elapsed:.1f} s, {threads} threads",
    }


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's)")
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--workload", default="layered_768")
    parser.add_argument("--rows", type=int, default=0, help="fused-kernel tile rows (0 = default)")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    parser.add_argument("--cpu-seconds", type=float, default=10.0)
    parser.add_argument("--intensity-from", type=int, default=None,
                        help="step at which |E|² accumulation starts (-1: never; default: the workload's)")
    args = parser.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU slab decomposition is "
                                   "design-only (see DESIGN.md)", "n_gpus": args.gpus}))
        return 2

    from diaphane_amd import workloads
    from diaphane_amd.gpu import Simulation, headless_context

    work = workloads.WORKLOADS[args.workload]()
    steps = args.steps if args.steps is not None else work.steps
    accumulate_from = work.accumulate_from
    if accumulate_from is not None and args.steps is not None:
        accumulate_from = steps // 2
    if args.intensity_from is not None:
        accumulate_from = None if args.intensity_from < 0 else min(args.intensity_from, steps)

    context = headless_context()
    sim = Simulation(context, work.scene)
    if args.rows:
        sim.set_tile(args.rows)
    cells = sim.cells

    # The step's access mix streamed with no stencil: the practical ceiling.
    probe_gbs = sim.stream_probe(5)

    if args.warmup:
        sim.advance_by(args.warmup)
        sim.wait()
        sim.reset()
        sim.wait()

    # Timed region: the workload's own step sequence from step 0.
    phase_a = accumulate_from if accumulate_from is not None else steps
    phase_b = steps - phase_a
    sim.wait()
    t0 = time.perf_counter()
    sim.record_event(0)
    sim.advance_by(phase_a)
    sim.record_event(1)
    if phase_b:
        sim.set_accumulate_intensity(True)
        sim.record_event(2)
        sim.advance_by(phase_b)
    sim.record_event(3)
    sim.wait()
    t1 = time.perf_counter()
    wall = t1 - t0

    ms_a = sim.elapsed_ms(0, 1)
    ms_b = sim.elapsed_ms(2, 3) if phase_b else 0.0
    bytes_a = BYTES_PER_CELL * cells * phase_a
    bytes_b = BYTES_PER_CELL_INTENSITY * cells * phase_b
    kernel_s = (ms_a + ms_b) / 1e3
    achieved = (bytes_a + bytes_b) / kernel_s / 1e9
    value = cells * steps / wall

    traffic = None
    traffic_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                f"k_fused_aligned_{work.name}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as handle:
            measured = json.load(handle)
        traffic = measured.get("traffic_bytes_per_launch")

    result = {
        "metric": "Yee cell-updates/s & %HBM roofline",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": work.name,
            "cells": cells,
            "extent": list(sim.extent),
            "steps_with_intensity": phase_b,
            "note": work.note,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_fused_aligned",
            "stream_probe_gbs": probe_gbs,
            "frac_of_stream_probe": achieved / probe_gbs,
            "bytes_per_cell": {"plain": BYTES_PER_CELL, "with_intensity_f64": BYTES_PER_CELL_INTENSITY},
            "avg_launch_ms": {"plain": ms_a / max(phase_a, 1), "with_intensity": ms_b / max(phase_b, 1)},
            "phase_gbs": {
                "plain": bytes_a / (ms_a / 1e3) / 1e9 if ms_a else None,
                "with_intensity": bytes_b / (ms_b / 1e3) / 1e9 if ms_b else None,
            },
        },
    }
    sim.close()
    if not args.no_cpu:
        cpu = cpu_baseline(args.workload, args.cpu_seconds)
        result["cpu_baseline"] = cpu
        result["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
