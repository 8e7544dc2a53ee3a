"""The unit tests of src/cpu.rs (:774-1035), ported onto the oracle. They pin
the stepping, energy, reversal, reset and refusal behaviour the oracle
restates; the device is held to the oracle bit for bit elsewhere."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Grid, Material, Scene, Shape, Source, Waveform


def pulse_scene(extent: Extent) -> Scene:
    # src/cpu.rs:765-773
    grid = Grid.new(extent, 1e-3)
    frequency = grid.frequency_for_resolution(20.0)
    return Scene(grid).with_source(Source.point([0.0, 0.0, 0.0], Axis.Z, Waveform.ricker(frequency)))


def fields(sim) -> np.ndarray:
    return np.concatenate([sim.electric_all, sim.magnetic_all]).copy()


def test_starts_at_rest():
    sim = oracle.Simulation(pulse_scene(Extent.cube(24)))
    assert sim.step_count() == 0
    assert sim.time() == 0.0
    assert sim.energy().total() == 0.0


def test_a_source_puts_energy_in_and_the_absorber_takes_it_out():
    sim = oracle.Simulation(pulse_scene(Extent.cube(48)))
    sim.advance_by(120)
    peak = sim.energy().total()
    assert peak > 0.0, "the source did nothing"
    assert sim.is_finite()
    sim.advance_by(600)
    remaining = sim.energy().total()
    assert remaining < peak * 1e-4, f"{remaining:e} left of {peak:e}: the absorber is leaking"


def mean_energy(sim, steps: int) -> float:
    total = 0.0
    for _ in range(steps):
        sim.advance()
        total += sim.energy().total()
    return total / steps


def test_a_pec_box_holds_on_to_its_energy():
    sim = oracle.Simulation(pulse_scene(Extent.cube(40)).with_boundary(Boundary.PEC))
    sim.advance_by(200)
    settled = mean_energy(sim, 200)
    sim.advance_by(2000)
    later = mean_energy(sim, 200)
    drift = abs(later - settled) / settled
    assert drift < 1e-3, f"energy drifted by {drift:e}"


def test_the_two_energies_alternate_in_a_cavity():
    sim = oracle.Simulation(Scene.cavity(Extent.cube(40)))
    sim.advance_by(400)
    lo, hi = np.inf, -np.inf
    for _ in range(200):
        sim.advance()
        energy = sim.energy()
        split = (energy.electric - energy.magnetic) / energy.total()
        lo, hi = min(lo, split), max(hi, split)
    assert hi - lo > 0.3, f"the split barely moved: {lo} to {hi}"


def test_a_pulse_expands_at_about_the_speed_of_light():
    extent = Extent.cube(80)
    sim = oracle.Simulation(pulse_scene(extent))
    centre, radius = 40, 25
    probe = [centre + radius, centre, centre]
    arrival = None
    for step in range(1, 400):
        sim.advance()
        if abs(sim.sample_electric(Axis.Z, probe)) > 1e-4:
            arrival = step
            break
    assert arrival is not None, "the pulse never reached the probe"
    travel = radius / sim.grid.courant
    assert arrival > travel * 0.9, f"arrived at step {arrival}, faster than light ({travel} steps)"
    assert arrival < travel * 3.0, f"arrived at step {arrival}, far too slow"


def test_a_perfect_conductor_expels_the_field():
    scene = pulse_scene(Extent.cube(40))
    metal = scene.materials.push(Material.PERFECT_CONDUCTOR)
    # cells 26..34 along x are 6..14 mm
    scene.shapes.append(Shape.Block([10e-3, 0.0, 0.0], [8e-3, 20e-3, 20e-3], metal))
    sim = oracle.Simulation(scene)
    sim.advance_by(300)
    ez = sim.electric(Axis.Z).reshape(40, 40, 40)  # [z, y, x]
    assert np.any(ez != 0.0), "the pulse never started"
    block = ez[10:30, 10:30, 26:34]
    assert np.all(block == 0.0), f"field leaked into the conductor: {np.abs(block).max():e}"


def test_stepping_back_undoes_stepping_forward():
    sim = oracle.Simulation(pulse_scene(Extent.cube(32)).with_boundary(Boundary.PEC))
    assert sim.is_reversible()
    sim.advance_by(60)
    before = fields(sim)
    peak = float(np.abs(before).max())
    assert peak > 0.0
    sim.advance_by(200)
    sim.reverse_by(200)
    assert sim.step_count() == 60
    worst = float(np.abs(fields(sim) - before).max())
    assert worst < 1e-4 * peak, f"round trip drifted by {worst / peak:e}"


def test_reversing_all_the_way_returns_to_an_empty_domain():
    sim = oracle.Simulation(pulse_scene(Extent.cube(24)).with_boundary(Boundary.PEC))
    sim.advance_by(120)
    excited = sim.energy().total()
    sim.reverse_by(120)
    assert sim.step_count() == 0
    residue = sim.energy().total()
    assert residue < excited * 1e-8, f"{residue:e} left of {excited:e} after unwinding to t = 0"


def test_reversal_refuses_a_scene_with_an_absorber():
    sim = oracle.Simulation(pulse_scene(Extent.cube(32)))
    assert not sim.is_reversible()
    sim.advance()
    with pytest.raises(oracle.OracleError, match="lossy"):
        sim.reverse_by(1)


def test_reversal_refuses_a_scene_with_a_conductor():
    scene = pulse_scene(Extent.cube(32)).with_boundary(Boundary.PEC)
    metal = scene.materials.push(Material.PERFECT_CONDUCTOR)
    scene.shapes.append(Shape.Block([8e-3, 0.0, 0.0], [4e-3] * 3, metal))
    sim = oracle.Simulation(scene)
    assert not sim.is_reversible()
    sim.advance()
    with pytest.raises(oracle.OracleError, match="lossy"):
        sim.reverse_by(1)


def test_reversal_refuses_to_pass_the_start():
    sim = oracle.Simulation(pulse_scene(Extent.cube(24)).with_boundary(Boundary.PEC))
    with pytest.raises(oracle.OracleError, match="cannot step back past the start"):
        sim.reverse_by(1)


def test_reset_returns_to_the_initial_state():
    sim = oracle.Simulation(pulse_scene(Extent.cube(24)))
    sim.advance_by(50)
    assert sim.energy().total() > 0.0
    sim.reset()
    assert sim.step_count() == 0
    assert sim.energy().total() == 0.0
    assert not np.any(fields(sim))


def energy_density(sim) -> np.ndarray:
    # cpu::Simulation::energy_density (src/cpu.rs:691-703): per cell,
    # Σ_axis ½(e² + h²) in f32, with no material or volume factors.
    density = np.zeros(sim.cells, dtype=np.float32)
    half = np.float32(0.5)
    for axis in range(3):
        e = sim.electric(axis)
        h = sim.magnetic(axis)
        density = density + half * (e * e + h * h)
    return density


def test_energy_density_sums_to_the_total_energy():
    sim = oracle.Simulation(pulse_scene(Extent.cube(24)))
    sim.advance_by(60)
    nx, ny, nz = sim.extent
    widths = [sim.grid.primary(Axis(a)).astype(np.float32) for a in range(3)]
    volume = (widths[2][:, None, None] * widths[1][None, :, None] * widths[0][None, None, :]).astype(np.float64)
    summed = float(np.sum(volume.ravel() * energy_density(sim).astype(np.float64)))
    total = sim.energy().total()
    assert abs(summed - total) < total * 1e-4, f"{summed} vs {total}"
