#!/usr/bin/env python3
"""Benchmark: Yee cell-updates/s of the fused gfx950 step, against the HBM
roofline, with the CPU oracle timed beside it.

Default workload (``layered_768``): a 768³ uniform grid (452,984,832 cells),
40 random dielectric z-slabs indexed through u8 into 16 materials, a centred
z-dipole, 1000 steps with |E|² accumulated (f64) over the last 500. Fields and
materials are resident in HBM before the timed region starts.

One step = one full leapfrog step of the whole grid (H, injections, E,
injections, intensity). Prints one JSON line. The default line also carries
``batched``, ``served`` (50 req/s) and ``served_200`` (200 req/s) objects: the
two workloads below, each against the same work run one solve at a time on
the same GPU (``--no-batched``, ``--no-served`` leave them out),
``other_workloads`` the bench lines of ``free_space_80`` and
``graded_scattering_256`` (each run of this script with ``--workload``, in a
child process; ``--no-other-workloads`` leaves them out), and
``cpu_baseline`` times the C oracle on all the run's host threads and on one.

``--workload batched_uneven_256``: 256 independent uneven solves (sides
log-uniform in [48, 160] per axis, PEC or absorbing, seeded cavity modes),
1000 steps each, advanced one launch per two steps for the whole batch; the same
solves run one at a time on the same GPU are timed beside it, and a sample of
members is checked bit for bit between the two.

``--workload served_stream``: Poisson arrivals (``--rate`` req/s) of solve
requests of mixed kinds, sizes and step counts, served by the continuous-
batching request loop under a 200 GiB device-memory budget, then the same
stream served one request at a time; reports aggregate cell-updates/s and
p50/p99 completion latency for both.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
BYTES_PER_CELL = 49            # 6 f32 fields read+written + 1 u8 material index
BYTES_PER_CELL_INTENSITY = 65  # + f64 intensity read-modify-write


def time_oracle(scene, threads: int, seconds: float) -> tuple:
    """Steps the C oracle on `threads` host threads until `seconds` have
    passed (at least 3 steps, after 2 warm-up): (cell-updates/s, steps, s)."""
    import oracle

    sim = oracle.Simulation(scene, threads=threads)
    sim.advance_by(2)  # warm-up
    steps = 0
    start = time.perf_counter()
    while True:
        sim.advance_by(1)
        steps += 1
        elapsed = time.perf_counter() - start
        if elapsed >= seconds and steps >= 3:
            break
    return sim.cells * steps / elapsed, steps, elapsed


def cpu_baseline(workload_name: str, seconds: float = 10.0) -> dict:
    """Times the C oracle (a restatement of the reference's cpu.rs) on the
    host's cores over a bounded sample of the same workload: once on every
    core the run may use (OpenMP over z, bit-identical to one thread), and
    once on a single thread, the reference's own timing discipline."""
    from diaphane_amd import workloads

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    if workload_name == "layered_768":
        sample = workloads.layered_768(nz=64)
        label = "768x768x64 slab of the layered_768 stack"
    else:
        sample = workloads.WORKLOADS[workload_name]()
        label = f"{workload_name} (full grid)"
    rate, steps, elapsed = time_oracle(sample.scene, threads, seconds)
    one, one_steps, one_elapsed = time_oracle(sample.scene, 1, seconds)
    return {
        "value": rate,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{label}: {steps} steps after 2 warm-up, {elapsed:.1f} s, {threads} threads",
        "single_thread": {
            "value": one,
            "unit": "cell-updates/s",
            "cores": 1,
            "sample": f"{label}: {one_steps} steps after 2 warm-up, {one_elapsed:.1f} s, 1 thread",
        },
    }


def log(message: str) -> None:
    print(f"[bench] {message}", file=sys.stderr, flush=True)


def batch_cpu_baseline(instances, seconds: float) -> dict:
    """The C oracle over whole instances of the batch, one after another, until
    the time budget is spent."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    updates = 0
    done = 0
    start = time.perf_counter()
    for instance in instances:
        sim = oracle.Simulation(instance.scene, threads=threads)
        e, h = instance.initial_fields()
        sim.restore_fields(e, h, 0)
        t = time.perf_counter()
        steps = 0
        while steps < 50:
            sim.advance_by(5)
            steps += 5
        updates += sim.cells * steps
        done += 1
        if time.perf_counter() - start >= seconds:
            break
    elapsed = time.perf_counter() - start
    return {
        "value": updates / elapsed,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {done} instances, 50 steps each, one after another, {elapsed:.1f} s "
                  f"(includes per-instance setup), {threads} threads",
    }


def run_batched(context, steps: int, warmup: int, tile=None, members: str = "all", sequential: bool = True) -> dict:
    """The 256 uneven solves run one at a time, then as one batch from the
    same initial fields; eight members are compared bit for bit between the
    two runs. Returns the measurements (tile: (rows, depth, flags) or None
    for the autotuner; members: "all", or only the "pec" cavities or the
    "absorbing" boxes; sequential False skips the one-at-a-time run)."""
    import numpy as np

    from diaphane_amd import workloads
    from diaphane_amd.batch import Batch
    from diaphane_amd.gpu import Simulation

    instances = [i for i in workloads.batched_uneven_256() if members in ("all", i.scene.boundary.kind)]
    log(f"creating {len(instances)} sims")
    sims = []
    initial = []
    for k, instance in enumerate(instances):
        sim = Simulation(context, instance.scene)
        e, h = instance.initial_fields()
        sim.write_fields(e, h, 0)
        sims.append(sim)
        initial.append((e, h))
        if (k + 1) % 64 == 0:
            log(f"  {k + 1} created")
    cells = [s.cells for s in sims]
    total_cells = sum(cells)

    # One at a time: each solve runs its steps alone, then the next.
    log("one at a time" if sequential else "one at a time: skipped")
    if warmup and sequential:
        sims[0].advance_by(warmup)
        sims[0].wait()
        sims[0].write_fields(*initial[0], 0)
    seq_kernel_ms = 0.0
    t0 = time.perf_counter()
    for sim in sims if sequential else []:
        sim.record_event(0)
        sim.advance_by(steps)
        sim.record_event(1)
        sim.wait()
        seq_kernel_ms += sim.elapsed_ms(0, 1)
    seq_wall = time.perf_counter() - t0
    check = list(range(0, len(sims), max(1, len(sims) // 8))) if sequential else []
    solo_fields = {k: (sims[k].read_electric(), sims[k].read_magnetic()) for k in check}

    log("batched")
    for sim, (e, h) in zip(sims, initial):
        sim.write_fields(e, h, 0)
    batch = Batch(context)
    for sim in sims:
        batch.add(sim, steps)
    if tile:
        batch.set_tile(*tile)
        tuned = dict(batch.tiling(), ms_per_step=None)
    else:
        rows, depth, ms = batch.autotune(3)
        tuned = dict(batch.tiling(), ms_per_step=ms)
    log(f"tiling {tuned}")
    batch.wait()
    t0 = time.perf_counter()
    batch.record_event(0)
    batch.advance_by(steps)
    batch.record_event(1)
    batch.wait()
    batch_wall = time.perf_counter() - t0
    batch_ms = batch.elapsed_ms(0, 1)
    assert all(s.step_count() == steps for s in sims)
    identical = all(
        np.array_equal(sims[k].read_electric().view(np.uint32), solo_fields[k][0].view(np.uint32))
        and np.array_equal(sims[k].read_magnetic().view(np.uint32), solo_fields[k][1].view(np.uint32))
        for k in check)
    launches = batch.launch_count()
    batch.close()
    for sim in sims:
        sim.close()

    updates = total_cells * steps
    achieved = BYTES_PER_CELL * updates / (batch_ms / 1e3) / 1e9
    log(f"  batched {updates / batch_wall / 1e9:.1f} G/s" +
        (f", one at a time {updates / seq_wall / 1e9:.1f} G/s" if sequential else ""))
    return {
        "instances": instances,
        "value": updates / batch_wall,
        "wall_s": batch_wall,
        "steps": steps,
        "cells": total_cells,
        "cells_min_max": [min(cells), max(cells)],
        "pec_instances": sum(1 for i in instances if i.scene.boundary.kind == "pec"),
        "sequential": {
            "value": updates / seq_wall,
            "unit": "cell-updates/s",
            "wall_s": seq_wall,
            "kernel_s": seq_kernel_ms / 1e3,
            "note": "the same solves from the same initial fields, dp_sim_advance one solve after another",
        } if sequential else None,
        "speedup_vs_sequential": seq_wall / batch_wall if sequential else None,
        "members_bit_identical_to_solo": identical if sequential else None,
        "members_checked": len(check),
        "tiling": tuned,
        "kernel": ("k_fused_stack_batch" if tuned.get("stack_waves") else "k_fused_pair_batch")
                  if tuned.get("two_step") else
                  ("k_fused_batch_direct" if tuned.get("wave_per_tile") else "k_fused_batch"),
        "launches": launches,
        "kernel_ms": batch_ms,
        "achieved_gbs": achieved,
    }


def batch_traffic(run: dict):
    """HBM bytes per launch of the batch kernel on the 256 grids, from its
    committed rocprofv3 counter summary when the run used the shape that
    summary measured (two steps per launch, 8 rows, 32-plane slabs, all
    members); None otherwise."""
    tiling = run["tiling"]
    if not (tiling.get("two_step") and not tiling.get("stack_waves") and tiling.get("rows") == 8
            and tiling.get("depth") == 32 and len(run["instances"]) == 256):
        return None, None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "k_fused_pair_batch_batched_uneven_256.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as handle:
        traffic = json.load(handle).get("traffic_bytes_per_launch")
    return traffic, (f"profiles/{os.path.basename(path)}: rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE per launch of this "
                     f"kernel at this shape, from a separate counter run (not measured in this run)")


def other_workload(name: str, warmup: int) -> dict:
    """The bench line of another single-grid workload, from a run of this
    script with --workload in a child process, summarised."""
    import subprocess

    cmd = [sys.executable, os.path.abspath(__file__), "--workload", name, "--no-cpu", "--warmup", str(warmup)]
    log(f"other workload: {name}")
    try:
        run = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return {"workload": name, "error": "timed out after 600 s"}
    lines = [line for line in run.stdout.splitlines() if line.startswith("{")]
    if run.returncode != 0 or not lines:
        return {"workload": name, "error": f"exit status {run.returncode}", "stderr_tail": run.stderr[-2000:]}
    d = json.loads(lines[-1])
    config = d["config"]
    roofline = d["roofline"]
    return {
        "workload": name,
        "metric": d["metric"],
        "value": d["value"],
        "unit": d["unit"],
        "steps": d["steps"],
        "ms_per_step": d["ms_per_step"],
        "cells": config["cells"],
        "extent": config["extent"],
        "readback_every": config.get("readback_every"),
        "tiling": config.get("tiling"),
        "roofline": {key: roofline.get(key) for key in ("bound", "achieved", "peak", "unit", "frac", "kernel",
                                                          "avg_launch_ms")},
        "note": config.get("note"),
    }


def batched_summary(run: dict) -> dict:
    """The batched form's sub-object of the default benchmark line."""
    traffic, source = batch_traffic(run)
    return {
        "workload": "batched_uneven_256",
        "metric": "aggregate Yee cell-updates/s, batched vs one at a time",
        "unit": "cell-updates/s",
        "value": run["value"],
        "sequential_value": run["sequential"]["value"],
        "speedup_vs_sequential": run["speedup_vs_sequential"],
        "members_bit_identical_to_solo": run["members_bit_identical_to_solo"],
        "members_checked": run["members_checked"],
        "instances": len(run["instances"]),
        "cells": run["cells"],
        "steps": run["steps"],
        "wall_s": run["wall_s"],
        "sequential_wall_s": run["sequential"]["wall_s"],
        "kernel": run["kernel"],
        "tiling": run["tiling"],
        "achieved_gbs": run["achieved_gbs"],
        "frac_of_hbm_peak": run["achieved_gbs"] / HBM_PEAK_GBS,
        "traffic_bytes_per_launch": traffic,
        "traffic_source": source,
    }


def bench_batched(args) -> int:
    """Aggregate cell-updates/s of the batched form against the same solves
    run one at a time on the same GPU."""
    from diaphane_amd.gpu import headless_context

    steps = args.steps if args.steps is not None else 1000
    tile = (args.rows, args.depth, args.tile_flags) if args.rows else None
    run = run_batched(headless_context(), steps, args.warmup, tile, args.members, not args.no_sequential)
    value = run["value"]
    result = {
        "metric": "aggregate Yee cell-updates/s, batched vs one at a time",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": run["wall_s"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched_uneven_256",
            "instances": len(run["instances"]),
            "cells": run["cells"],
            "cells_min_max": run["cells_min_max"],
            "pec_instances": run["pec_instances"],
            "members": args.members,
        },
        "sequential": run["sequential"],
        "speedup_vs_sequential": run["speedup_vs_sequential"],
        "members_bit_identical_to_solo": run["members_bit_identical_to_solo"],
        "members_checked": run["members_checked"],
        "autotune": run["tiling"],
        "roofline": {
            "bound": "hbm",
            "achieved": run["achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": run["achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": batch_traffic(run)[0],
            "traffic_source": batch_traffic(run)[1],
            "kernel": run["kernel"],
            "bytes_per_cell": BYTES_PER_CELL,
            "launches": run["launches"],
            "avg_launch_ms": run["kernel_ms"] / max(run["launches"], 1),
        },
    }
    if not args.no_cpu:
        cpu = batch_cpu_baseline(run["instances"], args.cpu_seconds)
        result["cpu_baseline"] = cpu
        result["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(result), flush=True)
    return 0


def serve_stream(context, stream, one_at_a_time: bool, chunk_steps: int, rows: int, depth: int,
                 chunk_cells: int = 0) -> dict:
    """Feeds the stream to a server at its arrival times; returns per-request
    latencies and the aggregate rate."""
    from diaphane_amd.serve import DONE, Server

    server = Server(context, one_at_a_time=one_at_a_time, chunk_steps=chunk_steps, rows=rows, depth=depth,
                    host_threads=4, chunk_cells=chunk_cells)
    pending = {}
    submitted_at = {}
    infos = []
    host_latency = []
    peak_bytes = 0
    start = time.perf_counter()
    next_log = start + 10.0
    i = 0
    while i < len(stream) or pending:
        now = time.perf_counter() - start
        while i < len(stream) and stream[i].arrival_s <= now:
            ticket = server.submit(stream[i].scene, stream[i].steps)
            pending[ticket] = i
            submitted_at[ticket] = time.perf_counter()
            i += 1
        peak_bytes = max(peak_bytes, server.stats()["used_bytes"])
        for ticket in list(pending):
            info = server.info(ticket)
            if info.state == DONE:
                server.wait(ticket)  # raises if the request failed
                server.result(ticket).read_electric()  # the result on the host
                host_latency.append(time.perf_counter() - submitted_at.pop(ticket))
                infos.append((pending.pop(ticket), info))
                server.release(ticket)
        if time.perf_counter() >= next_log:
            log(f"  {len(infos)}/{len(stream)} done, {len(pending)} in flight")
            next_log += 10.0
        wait = 0.001 if i >= len(stream) else min(0.001, max(0.0, stream[i].arrival_s - now))
        time.sleep(wait)
    stats = server.stats()
    server.close()
    infos.sort()
    latencies = np.array([info.latency_s for _, info in infos])
    on_host = np.array(host_latency)
    first = min(info.submitted_s for _, info in infos)
    last = max(info.completed_s for _, info in infos)
    updates = sum(info.cells * info.steps for _, info in infos)
    return {
        "value": updates / (last - first),
        "unit": "cell-updates/s",
        "makespan_s": last - first,
        "latency_s": {"p50": float(np.percentile(on_host, 50)), "p99": float(np.percentile(on_host, 99)),
                      "mean": float(on_host.mean()), "max": float(on_host.max()),
                      "note": "submission to the E field read back on the host"},
        "device_completion_s": {"p50": float(np.percentile(latencies, 50)),
                                "p99": float(np.percentile(latencies, 99))},
        "peak_resident_bytes": peak_bytes,
        "chunks": stats["chunks"],
        "device_busy_s": stats["busy_s"],
        "kernel_s": stats["kernel_s"],
        "kernel_rate": stats["cell_updates"] / stats["kernel_s"] if stats["kernel_s"] else None,
        "device_utilization": stats["kernel_s"] / (last - first),
    }


def run_served(context, requests: int, rate: float, chunk_steps: int, rows: int, chunk_cells: int) -> dict:
    """The Poisson request stream served with continuous batching, then the
    same stream (same arrival times) served one request at a time."""
    from diaphane_amd import workloads

    stream = workloads.served_stream(requests, rate)
    updates = sum(r.scene.grid.extent.total() * r.steps for r in stream)
    log(f"{len(stream)} requests over {stream[-1].arrival_s:.1f} s, {updates / 1e9:.0f} G cell-updates")
    # warm the code paths (kernels load on first launch)
    warm = workloads.served_stream(3, 1000.0, seed=99, lo=32, hi=40, steps_lo=40, steps_hi=40)
    cells_per_chunk = (1 << 64) - 1 if chunk_cells < 0 else chunk_cells
    serve_stream(context, warm, False, chunk_steps, rows, 0, cells_per_chunk)
    serve_stream(context, warm, True, chunk_steps, rows, 0)
    log("continuous batching")
    batched = serve_stream(context, stream, False, chunk_steps, rows, 0, cells_per_chunk)
    log(f"  {batched['value'] / 1e9:.1f} G/s, p50 {batched['latency_s']['p50']:.2f} s")
    log("one at a time")
    sequential = serve_stream(context, stream, True, chunk_steps, rows, 0)
    log(f"  {sequential['value'] / 1e9:.1f} G/s, p50 {sequential['latency_s']['p50']:.2f} s")
    return {
        "stream": stream,
        "cell_updates": updates,
        "served": batched,
        "sequential": sequential,
        "speedup_vs_sequential": batched["value"] / sequential["value"],
        "kinds": {k: sum(1 for r in stream if r.kind == k) for k in workloads.SERVED_KINDS},
        "chunk_cells": "every member" if chunk_cells < 0 else (chunk_cells or 128 << 20),
    }


def served_summary(run: dict, rate: float, chunk_steps: int) -> dict:
    """The served form's sub-object of the default benchmark line."""
    return {
        "workload": "served_stream",
        "metric": "served aggregate Yee cell-updates/s with p50/p99 completion latency",
        "unit": "cell-updates/s",
        "value": run["served"]["value"],
        "latency_s": run["served"]["latency_s"],
        "sequential_value": run["sequential"]["value"],
        "sequential_latency_s": run["sequential"]["latency_s"],
        "speedup_vs_sequential": run["speedup_vs_sequential"],
        "requests": len(run["stream"]),
        "rate_per_s": rate,
        "cell_updates": run["cell_updates"],
        "kinds": run["kinds"],
        "hbm_budget_bytes": 200 << 30,
        "peak_resident_bytes": run["served"]["peak_resident_bytes"],
        "device_utilization": run["served"]["device_utilization"],
        "kernel_rate": run["served"]["kernel_rate"],
        "sequential_kernel_rate": run["sequential"]["kernel_rate"],
        "chunk_steps": chunk_steps,
        "chunk_cells": run["chunk_cells"],
        # No schedule can finish before the last request arrives: the stream's
        # cell-updates over its arrival span bound any served rate.
        "arrival_bound_value": run["cell_updates"] / max(run["stream"][-1].arrival_s - run["stream"][0].arrival_s,
                                                         1e-9),
    }


def bench_served(args) -> int:
    from diaphane_amd.gpu import headless_context

    run = run_served(headless_context(), args.requests, args.rate, args.chunk_steps, args.rows, args.chunk_cells)
    stream = run["stream"]
    result = {
        "metric": "served aggregate Yee cell-updates/s with p50/p99 completion latency",
        "value": run["served"]["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": int(sum(r.steps for r in stream)),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "served_stream",
            "requests": len(stream),
            "rate_per_s": args.rate,
            "cell_updates": run["cell_updates"],
            "kinds": run["kinds"],
            "hbm_budget_bytes": 200 << 30,
            "chunk_steps": args.chunk_steps,
            "chunk_cells": run["chunk_cells"],
        },
        "served": run["served"],
        "sequential": run["sequential"],
        "speedup_vs_sequential": run["speedup_vs_sequential"],
    }
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's)")
    parser.add_argument("--warmup", type=int, default=20)
    parser.add_argument("--workload", default="layered_768")
    parser.add_argument("--rows", type=int, default=0, help="fused-kernel tile rows (0 = default / autotune)")
    parser.add_argument("--depth", type=int, default=0, help="fused-kernel z-slab depth (0 = default)")
    parser.add_argument("--tile-flags", type=int, default=-1,
                        help="tile flags (1 XCD remap, run length in bits 8..15; 2 non-temporal stores; 4 buffer "
                             "addressing; 8 two steps per launch; 32 stacked waves, 2 or 4 in bits 16..19; "
                             "batched: 16 one wave per tile; -1 = default)")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU-oracle baseline")
    parser.add_argument("--no-batched", action="store_true",
                        help="default workload: skip the batched_uneven_256 comparison")
    parser.add_argument("--no-served", action="store_true",
                        help="default workload: skip the served_stream comparison")
    parser.add_argument("--no-other-workloads", action="store_true",
                        help="default workload: skip the free_space_80 and graded_scattering_256 lines")
    parser.add_argument("--cpu-seconds", type=float, default=10.0)
    parser.add_argument("--members", choices=("all", "pec", "absorbing"), default="all",
                        help="batched_uneven_256: all members, or only the PEC cavities or the absorbing boxes")
    parser.add_argument("--no-sequential", action="store_true",
                        help="batched_uneven_256: skip the one-at-a-time run")
    parser.add_argument("--requests", type=int, default=200, help="served_stream: number of requests")
    parser.add_argument("--rate", type=float, default=50.0, help="served_stream: arrivals per second")
    parser.add_argument("--chunk-steps", type=int, default=32, help="served_stream: launches between joins")
    parser.add_argument("--chunk-cells", type=int, default=0,
                        help="served_stream: cells stepping per chunk, least remaining work first "
                             "(0: the server's default; -1: every member every chunk)")
    parser.add_argument("--intensity-from", type=int, default=None,
                        help="step at which |E|² accumulation starts (-1: never; default: the workload's)")
    args = parser.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; splitting a grid into z-slabs across "
                                   "GPUs is not built", "n_gpus": args.gpus}))
        return 2

    if args.workload == "batched_uneven_256":
        return bench_batched(args)
    if args.workload == "served_stream":
        return bench_served(args)

    from diaphane_amd import workloads
    from diaphane_amd.gpu import BUFFER_ELECTRIC, TILE_STACK, TILE_TWO_STEP, Simulation, headless_context

    work = workloads.WORKLOADS[args.workload]()
    steps = args.steps if args.steps is not None else work.steps
    accumulate_from = work.accumulate_from
    if accumulate_from is not None and args.steps is not None:
        accumulate_from = steps // 2
    if args.intensity_from is not None:
        accumulate_from = None if args.intensity_from < 0 else min(args.intensity_from, steps)

    context = headless_context()
    sim = Simulation(context, work.scene)
    if args.rows or args.depth or args.tile_flags >= 0:
        sim.set_tile(args.rows, args.depth, args.tile_flags)
    cells = sim.cells

    # The step's access mix streamed with no stencil: the practical ceiling.
    probe_gbs = sim.stream_probe(5)
    # Tile shape by the engine's autotuner unless given (the fields are unchanged).
    if args.rows or args.depth or args.tile_flags >= 0:
        tiling = {"rows": args.rows or 4, "depth": args.depth or 32, "flags": args.tile_flags, "autotuned": False}
    else:
        tiling = dict(sim.autotune(8), autotuned=True)
        if accumulate_from is not None and accumulate_from < steps:
            sim.set_accumulate_intensity(True)  # tune the accumulating steps' shape too
            tiling["with_intensity"] = sim.autotune(8)
            sim.set_accumulate_intensity(False)
    # The plain steps' tiling run as a pure copy (same tiles, rows, halos and
    # stores, no stencil): the ceiling for that tiling's memory traffic.
    tile_gbs = sim.tile_probe(tiling["rows"] if tiling["rows"] in (4, 6, 8) else 0, tiling["depth"], 5)

    if args.warmup:
        # The warm-up steps take the timed sequence's phases (and allocate the
        # intensity sum's buffer where it has one, as its first accumulating
        # step would).
        if accumulate_from is not None and accumulate_from < steps:
            sim.advance_by(args.warmup - args.warmup // 2)
            sim.set_accumulate_intensity(True)
            sim.advance_by(args.warmup // 2)
            sim.set_accumulate_intensity(False)
        else:
            sim.advance_by(args.warmup)
        sim.wait()
        sim.reset()
        sim.wait()

    # Timed region: the workload's own step sequence from step 0, with its
    # readbacks of E to the host where it has them.
    readback = work.readback_every if args.steps is None else None
    phase_a = accumulate_from if accumulate_from is not None else steps
    phase_b = steps - phase_a

    # The application's readback buffers: two page-locked host buffers from
    # the engine, used in turn (a viewer shows one frame while the next one
    # lands).
    frames = []
    if readback:
        from diaphane_amd.gpu import PinnedBuffer

        frames = [PinnedBuffer(3 * cells), PinnedBuffer(3 * cells)]
        # Setup, like the warm-up steps: the sim's two staging buffers are
        # allocated at their first readback and the host buffers touched.
        for frame in frames:
            sim.begin_read(BUFFER_ELECTRIC, frame.array)
        sim.finish_read()

    def run(count: int) -> None:
        if not readback:
            sim.advance_by(count)
            return
        done = 0
        while done < count:
            chunk = min(readback, count - done)
            sim.advance_by(chunk)
            # This chunk's E to the host, overlapped with the next chunk's
            # steps; once it is queued, the previous chunk's E has landed in
            # the other buffer before the next chunk is queued.
            sim.begin_read(BUFFER_ELECTRIC, frames[(done // readback) % 2].array)
            sim.wait_reads(1)
            done += chunk

    launches = {"plain": 0, "with_intensity": 0}

    def timed_pass(with_readback: bool):
        nonlocal readback
        saved, readback = readback, (readback if with_readback else None)
        sim.wait()
        start = time.perf_counter()
        first = sim.launch_count()
        sim.record_event(0)
        run(phase_a)
        sim.record_event(1)
        middle = sim.launch_count()
        if phase_b:
            sim.set_accumulate_intensity(True)
            sim.record_event(2)
            run(phase_b)
        sim.record_event(3)
        sim.finish_read()  # every readback lands inside the timed region
        sim.wait()
        elapsed = time.perf_counter() - start
        launches["plain"], launches["with_intensity"] = middle - first, sim.launch_count() - middle
        readback = saved
        return elapsed, sim.elapsed_ms(0, 1), sim.elapsed_ms(2, 3) if phase_b else 0.0

    wall, ms_a, ms_b = timed_pass(True)
    if readback:
        # The kernels' own time, from the same sequence without the readbacks.
        sim.reset()
        sim.set_accumulate_intensity(False)
        _, ms_a, ms_b = timed_pass(False)
    bytes_a = BYTES_PER_CELL * cells * phase_a
    bytes_b = BYTES_PER_CELL_INTENSITY * cells * phase_b
    kernel_s = (ms_a + ms_b) / 1e3
    achieved = (bytes_a + bytes_b) / kernel_s / 1e9
    value = cells * steps / wall

    # The kernel each phase ran: two steps per launch when its tiling says so.
    def kernel_of(flags: int) -> str:
        if flags >= 0 and flags & TILE_TWO_STEP:
            return "k_fused_stack" if flags & TILE_STACK else "k_fused_pair"
        return "k_fused_aligned"

    plain_flags = tiling.get("flags", -1)
    kernels = {"plain": kernel_of(plain_flags),
               "with_intensity": kernel_of(tiling.get("with_intensity", {}).get("flags", plain_flags))}
    dominant = "with_intensity" if ms_b > ms_a else "plain"
    # Measured HBM bytes per launch of the dominant phase's kernel, from its
    # committed rocprofv3 counter summary for this workload.
    traffic = None
    suffix = "" if dominant == "plain" else "_intensity"
    traffic_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                f"{kernels[dominant]}_{work.name}{suffix}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as handle:
            traffic = json.load(handle).get("traffic_bytes_per_launch")

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
            "readback_every": readback,
            "readback": ("E of each chunk into one of two page-locked buffers by dp_sim_begin_read, overlapped "
                         "with the next chunk's steps; each chunk waits (dp_sim_wait_reads) for the previous "
                         "chunk's E to land before queueing the next; all land inside the timed region")
            if readback else None,
            "tiling": tiling,
            "note": work.note,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kernels[dominant],
            "kernels": kernels,
            "launches": dict(launches),
            "achieved_definition": "49 B (65 B with the f64 intensity sum) per cell-update x cell-updates "
                                   "/ kernel time; a two-step launch does two cell-updates per cell",
            "stream_probe_gbs": probe_gbs,
            "frac_of_stream_probe": achieved / probe_gbs,
            "tile_probe_gbs": tile_gbs,
            # the single-step tiling's copy pattern: comparable only when
            # the plain steps ran that kernel
            "plain_frac_of_tile_probe": ((bytes_a / (ms_a / 1e3) / 1e9) / tile_gbs
                                         if ms_a and kernels["plain"] == "k_fused_aligned" else None),
            "bytes_per_cell": {"plain": BYTES_PER_CELL, "with_intensity_f64": BYTES_PER_CELL_INTENSITY},
            "avg_launch_ms": {"plain": ms_a / max(launches["plain"], 1),
                              "with_intensity": ms_b / max(launches["with_intensity"], 1)},
            "phase_gbs": {
                "plain": bytes_a / (ms_a / 1e3) / 1e9 if ms_a else None,
                "with_intensity": bytes_b / (ms_b / 1e3) / 1e9 if ms_b else None,
            },
        },
    }
    if traffic is not None:
        result["roofline"]["traffic_source"] = (
            f"profiles/{os.path.basename(traffic_file)}: rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE per launch of "
            f"this kernel at this shape, from a separate counter run (not measured in this run)")
    sim.close()
    if work.name == "layered_768":
        # The batched and served forms on their own workloads, each against
        # the same work run one solve at a time on this GPU.
        if not args.no_batched:
            result["batched"] = batched_summary(run_batched(context, 1000, args.warmup))
        if not args.no_served:
            result["served"] = served_summary(run_served(context, 200, 50.0, 32, 0, 0), 50.0, 32)
            # the top of the stated arrival range, where the device, not the
            # arrivals, bounds the served rate
            result["served_200"] = served_summary(run_served(context, 200, 200.0, 32, 0, 0), 200.0, 32)
        if not args.no_other_workloads:
            result["other_workloads"] = {name: other_workload(name, args.warmup)
                                         for name in ("free_space_80", "graded_scattering_256")}
    if not args.no_cpu:
        cpu = cpu_baseline(args.workload, args.cpu_seconds)
        result["cpu_baseline"] = cpu
        result["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
