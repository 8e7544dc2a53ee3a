"""The workloads at their full sizes on the device.

Each is held to the oracle where the oracle finishes in seconds (bit for bit,
plus the 1e-5-of-peak bar) and otherwise to size-independent properties:
determinism, agreement between the two kernel paths, energy conservation in a
closed box, equipartition of a travelling packet. The largest case runs a grid
whose field components exceed 4 GiB, which takes the 64-bit addressing path.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import AXES, Axis, Boundary, Extent, Grid, Refinement, Scene, Source, Waveform
from diaphane_amd import workloads
from diaphane_amd.gpu import KERNEL_TWO_PASS, TILE_BUFFER, TILE_STACK, TILE_TWO_STEP, TILE_XCD_REMAP, Simulation

from conftest import assert_field_parity

pytestmark = pytest.mark.gpu


def bits_equal(a, b) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def assert_oracle_parity(name, reference, candidate, steps):
    assert reference.step_count() == candidate.step_count() == steps
    for label, ref_all, cand in (("E", reference.electric_all, candidate.read_electric()),
                                 ("H", reference.magnetic_all, candidate.read_magnetic())):
        refs = [ref_all[a * reference.cells:(a + 1) * reference.cells] for a in range(3)]
        cands = [candidate.component(cand, a) for a in AXES]
        assert_field_parity(f"{name} {label}", refs, cands, steps, 1e-5)
        assert bits_equal(ref_all, cand), f"{name} {label} differs from the oracle after {steps} steps"


def test_free_space_80_at_full_size(context):
    # 80³, 500 steps, a readback every 5 steps; fields against the oracle at
    # checkpoints and equipartition of the travelling packet.
    work = workloads.free_space_80()
    reference = oracle.Simulation(work.scene)
    device = Simulation(context, work.scene)
    ue, uh = [], []
    for step in range(5, work.steps + 1, 5):
        device.advance_by(5)
        electric = device.read_electric()  # the workload's readback cadence
        assert np.isfinite(electric).all()
        energy = device.energy()
        ue.append(energy.electric)
        uh.append(energy.magnetic)
        if step % 100 == 0:
            reference.advance_by(step - reference.step_count())
            assert_oracle_parity("free_space_80", reference, device, step)
    ue, uh = np.array(ue[40:]), np.array(uh[40:])  # steps 205-500
    split = abs(ue.mean() - uh.mean()) / (ue.mean() + uh.mean())
    assert split < 0.1, f"electric/magnetic energy split {split:.3f}"


def test_graded_scattering_256_at_full_size(context):
    work = workloads.graded_scattering_256()
    device = Simulation(context, work.scene)
    assert device.extent[0] >= 250
    window = 60
    reference = oracle.Simulation(work.scene)
    reference.advance_by(window)
    device.advance_by(window)
    assert_oracle_parity("graded_scattering_256", reference, device, window)
    # the whole run twice: the device result is deterministic
    device.advance_by(work.steps - window)
    first = device.energy()
    again = Simulation(context, work.scene)
    again.advance_by(work.steps)
    second = again.energy()
    assert np.isfinite(first.total()) and first.total() > 0.0
    assert (first.electric, first.magnetic) == (second.electric, second.magnetic)


def test_layered_stack_slab_matches_the_oracle(context):
    # the 768³ stack's 768×768×64 slab, intensity on for the second half
    work = workloads.layered_768(nz=64)
    reference = oracle.Simulation(work.scene)
    device = Simulation(context, work.scene)
    for sim in (reference, device):
        sim.advance_by(30)
        sim.set_accumulate_intensity(True)
        sim.advance_by(30)
    assert_oracle_parity("layered_768 slab", reference, device, 60)
    assert np.array_equal(reference.intensity_f64(), device.read_intensity_f64())


def test_layered_768_at_full_size(context):
    # 1000 steps, |E|² over the last 500: the fused path and the two-pass path
    # (the reference's dispatch shape) agree bit for bit at full size.
    work = workloads.layered_768()
    results = []
    for kernel in (None, KERNEL_TWO_PASS):
        sim = Simulation(context, work.scene)
        if kernel is not None:
            sim.set_kernel(kernel)
        sim.advance_by(work.accumulate_from)
        sim.set_accumulate_intensity(True)
        sim.advance_by(work.steps - work.accumulate_from)
        assert sim.accumulated_steps() == 500
        energy = sim.energy()
        results.append((energy.electric, energy.magnetic, sim.read_electric()[::9973].copy(),
                        float(sim.read_intensity_f64().sum())))
        sim.close()
    (ue0, uh0, e0, i0), (ue1, uh1, e1, i1) = results
    assert np.isfinite(ue0) and ue0 > 0.0 and i0 > 0.0
    assert (ue0, uh0) == (ue1, uh1)
    assert bits_equal(e0, e1)
    assert i0 == i1


def mean_energy(sim, steps=200) -> float:
    total = 0.0
    for _ in range(steps):
        sim.advance_by(1)
        total += sim.energy().total()
    return total / steps


def test_a_closed_box_conserves_energy_for_40000_steps(context):
    # tests/validation.rs:339-372 on the device, at its 40,000 steps; the
    # fields after 40,800 steps are the oracle's bit for bit.
    scene = Scene.cavity(Extent.cube(36))
    device = Simulation(context, scene)
    device.advance_by(400)
    early = mean_energy(device)
    device.advance_by(40_000)
    late = mean_energy(device)
    drift = abs(late - early) / early
    print(f"uniform 36^3 PEC cavity: energy drift over 40000 steps {drift:.3e} (bar 1e-3)", flush=True)
    assert drift < 1e-3
    reference = oracle.Simulation(scene)
    reference.advance_by(device.step_count())
    assert bits_equal(reference.electric_all, device.read_electric())
    assert bits_equal(reference.magnetic_all, device.read_magnetic())


def test_a_graded_closed_box_conserves_energy_for_20000_steps(context):
    # tests/graded.rs:315-350 on the device: a 24 mm cavity graded to 0.4 mm
    # cells in a band across x; the reference publishes a drift of 4.9e-4
    # (docs/architecture.md:312). The fields after the run are the oracle's
    # bit for bit, so the drift is the oracle's too.
    grid = Grid.graded([0.024] * 3, 1000.0, [Refinement.across(Axis.X, 0.0, 8e-3, 0.4e-3)])
    scene = Scene.cavity_on(grid)
    device = Simulation(context, scene)
    device.advance_by(400)
    early = mean_energy(device)
    device.advance_by(20_000)
    late = mean_energy(device)
    drift = abs(late - early) / early
    print(f"graded cavity {device.extent}: energy drift over 20000 steps {drift:.3e} "
          f"(reference 4.9e-4, bar 1e-3)", flush=True)
    assert drift < 1e-3
    reference = oracle.Simulation(scene)
    reference.advance_by(device.step_count())
    assert bits_equal(reference.electric_all, device.read_electric())
    assert bits_equal(reference.magnetic_all, device.read_magnetic())


@pytest.mark.parametrize("tile", [None, (4, 0, TILE_TWO_STEP | TILE_STACK | TILE_BUFFER | TILE_XCD_REMAP)],
                         ids=["default", "stacked"])
def test_the_largest_grid_takes_the_64_bit_path(context, tile):
    # 1024 × 1024 × 1025 cells: one field component is 4.3 GB (> 4 GiB), so
    # the fused kernels address with 64-bit offsets (buffer addressing, asked
    # for, does not fit and is not used). A dipole off-centre, three steps
    # (with stacked tiles: a pair and a single), every cell against the oracle.
    grid = Grid(Extent(1024, 1024, 1025), 1e-3)
    scene = Scene(grid, Boundary.PEC)
    frequency = grid.frequency_for_resolution(16.0)
    scene.sources.append(Source.point([0.2, -0.1, 0.3], Axis.Z, Waveform.Ricker(frequency, 0.0)))
    cells = 1024 * 1024 * 1025
    assert cells * 4 > 0xFFFFFFFF
    device = Simulation(context, scene)
    if tile:
        device.set_tile(*tile)
    device.advance_by(3)
    reference = oracle.Simulation(scene)
    reference.advance_by(3)
    assert bits_equal(reference.electric_all, device.read_electric())
    assert bits_equal(reference.magnetic_all, device.read_magnetic())
    assert float(np.abs(reference.electric_all).max()) > 0.0
