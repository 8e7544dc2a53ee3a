"""The f32 update's error against an f64 build of the same update at the
workloads' full sizes: the graded 256³ scattering solve over its 2000 steps
and the 768×768×64 slab of the layered stack over its 1000 steps. The oracle
steps the f32 fields and its f64 track (the same update in double on the same
tables and injections); the device is bit-identical to the f32 fields, so
their distance is the engine's error. Minutes of CPU time, so it is a script,
not a test:

    python tests/f64_error_full_size.py [graded|layered ...]
"""
from __future__ import annotations

import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from diaphane_amd import workloads  # noqa: E402


def error_of(scene, steps: int, threads: int) -> dict:
    sim = oracle.Simulation(scene, threads=threads)
    sim.f64_begin()
    start = time.perf_counter()
    sim.advance_by(steps)
    sim.f64_advance_by(steps)
    out = {"cells": sim.cells, "steps": steps, "seconds": time.perf_counter() - start}
    for name, f32, f64 in (("E", sim.electric_all, sim.electric_f64()), ("H", sim.magnetic_all, sim.magnetic_f64())):
        peak = float(np.max(np.abs(f64)))
        out[name] = float(np.max(np.abs(f32.astype(np.float64) - f64))) / peak
        out[name + "_peak"] = peak
    return out


def main(which) -> int:
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    cases = {
        "graded": lambda: (workloads.graded_scattering_256().scene, 2000, "graded_scattering_256 (full grid)"),
        "layered": lambda: (workloads.layered_768(nz=64).scene, 1000, "layered stack, 768x768x64 slab"),
        # the same slab while the dipole's pulse is still inside it (by 1000 steps the 64-plane
        # slab's absorbing walls have taken all but a remnant, whose peak is small)
        "layered300": lambda: (workloads.layered_768(nz=64).scene, 300, "layered stack, 768x768x64 slab"),
    }
    for key in which or list(cases):
        scene, steps, label = cases[key]()
        r = error_of(scene, steps, threads)
        both = max(r["E_peak"], r["H_peak"])
        print(f"{label}: {r['cells']} cells, {steps} steps: E {r['E']:.2e} of peak {r['E_peak']:.3e}, "
              f"H {r['H']:.2e} of peak {r['H_peak']:.3e}; of the larger peak: E {r['E'] * r['E_peak'] / both:.2e}, "
              f"H {r['H'] * r['H_peak'] / both:.2e} ({r['seconds']:.0f} s on {threads} threads)", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
