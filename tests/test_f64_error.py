"""The f32 update's error against the same update in f64, per workload.

The device fields are bit-identical to the f32 oracle (the GPU parity tests
prove it), so the f32 oracle's distance from the oracle's f64 track is the
engine's error against an f64 build of the same update. Each workload runs at
a size the CPU finishes in seconds; the measured max|Δ| over the f64 peak is
printed and held to 1e-4 (the reference's own GPU-vs-CPU bar).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import workloads

BOUND = 1e-4


def f64_error(scene, steps, initial=None):
    sim = oracle.Simulation(scene)
    if initial is not None:
        sim.restore_fields(*initial, 0)
    sim.f64_begin()
    sim.advance_by(steps)
    sim.f64_advance_by(steps)
    out = {}
    for name, f32, f64 in (("E", sim.electric_all, sim.electric_f64()), ("H", sim.magnetic_all, sim.magnetic_f64())):
        peak = float(np.max(np.abs(f64)))
        out[name] = (float(np.max(np.abs(f32.astype(np.float64) - f64))) / peak, peak)
    return out


def report(label, cells, steps, errors):
    print(f"\n[f64 error] {label}: {cells} cells, {steps} steps: "
          f"E {errors['E'][0]:.2e} of peak {errors['E'][1]:.3e}, H {errors['H'][0]:.2e} of peak {errors['H'][1]:.3e}")
    for name, (relative, peak) in errors.items():
        assert peak > 0.0
        assert relative < BOUND, f"{label} {name}: f32 error {relative:e} of the f64 peak"


def test_free_space_80_full_size():
    work = workloads.free_space_80()
    errors = f64_error(work.scene, work.steps)
    report("free_space_80 (full)", work.scene.grid.extent.total(), work.steps, errors)


def test_graded_scattering_reduced():
    work = workloads.graded_scattering_256(side=96)
    steps = 400
    errors = f64_error(work.scene, steps)
    report("graded_scattering (96-cell sample)", work.scene.grid.extent.total(), steps, errors)


def test_layered_stack_reduced():
    work = workloads.layered_768(side=96)
    steps = 400
    errors = f64_error(work.scene, steps)
    report("layered stack (96³ sample)", work.scene.grid.extent.total(), steps, errors)


@pytest.mark.parametrize("index", [3])
def test_batched_instance_full_size(index):
    instance = workloads.batched_uneven_256()[index]
    errors = f64_error(instance.scene, 1000, instance.initial_fields())
    report(f"batched_uneven_256 instance {index} (full)", instance.scene.grid.extent.total(), 1000, errors)


@pytest.mark.parametrize("kind", workloads.SERVED_KINDS)
def test_served_kinds(kind):
    request = next(r for r in workloads.served_stream(12, seed=31, lo=32, hi=48) if r.kind == kind)
    # the open scenes while their pulse is still inside the grid (a fully
    # absorbed remnant has no meaningful peak to compare against); the
    # closed cavity rings for the full 1000
    steps = {"cavity_ringdown": 1000, "tfsf_sphere": 200, "lossy_half_space": 150}[kind]
    errors = f64_error(request.scene, steps)
    report(f"served {kind} ({request.scene.grid.extent.as_array()})", request.scene.grid.extent.total(), steps,
           errors)
