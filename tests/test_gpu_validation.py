"""The reference's analytic acceptance tests (tests/validation.rs,
tests/graded.rs), run on the device itself.

The device is bit-identical to the oracle, which these same tests pin on the
CPU (test_oracle_validation.py). Here they run directly against the engine,
through the C ABI, at the reference's own tolerances: an exact discrete plane
wave in four directions at two resolutions and in dielectrics (with the
negative control that shows the check can fail), the plane-wave source's
one-way launch, the PEC box's mirror symmetry, the absorbing layer's
reflection, and on graded grids a reversal round trip, agreement with a
uniformly fine grid and the grading transition's reflection.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Grid, Scene, Source, Waveform
from diaphane_amd.gpu import Simulation
from test_oracle_validation import (CELL, DOMAIN, FINE, TRANSVERSE, PlaneWave, band, pec_box, resample,
                                    uniformly_fine)

pytestmark = pytest.mark.gpu


class DeviceView:
    """The device sim seen through the oracle's accessors that PlaneWave uses."""

    def __init__(self, sim: Simulation):
        self.sim = sim
        self.extent = sim.extent
        self._e = self._h = None

    def refresh(self) -> None:
        self._e, self._h = self.sim.read_electric(), self.sim.read_magnetic()

    def step_count(self) -> int:
        return self.sim.step_count()

    def electric(self, a) -> np.ndarray:
        return self.sim.component(self._e, Axis(int(a)))

    def magnetic(self, a) -> np.ndarray:
        return self.sim.component(self._h, Axis(int(a)))


def device_plane_wave(context, wave: PlaneWave, scene: Scene, steps: int = 30) -> float:
    sim = Simulation(context, scene)
    e, h = wave.fields(sim.extent, 0.0)
    sim.write_fields(np.concatenate(e).astype(np.float32), np.concatenate(h).astype(np.float32), 0)
    sim.advance_by(steps)
    view = DeviceView(sim)
    view.refresh()
    error = wave.worst_error(view, steps + 2)
    sim.close()
    return error


@pytest.mark.parametrize("direction", [[1, 0, 0], [0, 0, 1], [1, 1, 1], [2, 1, 0]])
@pytest.mark.parametrize("cells", [10.0, 20.0])
def test_an_exact_plane_wave_propagates_exactly_on_the_device(context, direction, cells):
    # tests/validation.rs:189-240: within 1e-4 of the peak after 30 steps
    scene = pec_box(96)
    wave = PlaneWave(scene.grid.courant, direction, cells, 1.0)
    error = device_plane_wave(context, wave, scene)
    peak = float(max(np.abs(wave.h_amp).max(), np.abs(wave.e_amp).max()))
    assert error < 1e-4 * peak, f"{direction} at {cells} cells/λ: error {error:e}"


@pytest.mark.parametrize("permittivity", [2.25, 9.0])
def test_a_dielectric_plane_wave_propagates_exactly_on_the_device(context, permittivity):
    # tests/validation.rs:295-310
    index = math.sqrt(permittivity)
    scene = pec_box(96, permittivity)
    wave = PlaneWave(scene.grid.courant, [1, 0, 0], 20.0 * index, index)
    error = device_plane_wave(context, wave, scene)
    peak = float(max(np.abs(wave.h_amp).max(), np.abs(wave.e_amp).max()))
    assert error < 1e-4 * peak
    assert abs(wave.phase_velocity() - 1.0 / index) < 0.005 / index


def test_the_device_check_would_notice_a_wrong_dispersion_relation(context):
    # tests/validation.rs:312-337: the continuum phase rate must fail
    scene = pec_box(96)
    wave = PlaneWave(scene.grid.courant, [1, 0, 0], 10.0, 1.0)
    wave.phase_rate = scene.grid.courant * math.sqrt(float((wave.k * wave.k).sum()))
    assert device_plane_wave(context, wave, scene) > 0.05


def test_a_plane_wave_source_is_one_way_on_the_device(context):
    # tests/validation.rs:374-429: the backward leak stays below −25 dB
    scene = Scene.empty(Extent(120, 32, 32), 1e-3)
    frequency = scene.grid.frequency_for_resolution(20.0)
    scene.sources.append(Source.plane_wave(Axis.X, -0.020, Axis.Z, Waveform.gaussian_pulse(frequency, 4.0)))
    scene.validate()
    sim = Simulation(context, scene)
    plane = 40
    total_peak = scattered_peak = 0.0
    for _ in range(900):
        sim.advance_by(1)
        ez = sim.component(sim.read_electric(), Axis.Z).reshape(32, 32, 120)
        scattered_peak = max(scattered_peak, float(np.abs(ez[16, 16, 12:plane - 10]).max()))
        total_peak = max(total_peak, float(np.abs(ez[16, 16, plane + 10:108]).max()))
    sim.close()
    assert 0.8 <= total_peak < 1.25
    leak = 20.0 * math.log10(scattered_peak / total_peak)
    print(f"TFSF backward leak on the device: {leak:.1f} dB (bar -25 dB)", flush=True)
    assert leak < -25.0


def test_a_conducting_box_cannot_tell_its_low_walls_from_its_high_ones_on_the_device(context):
    # tests/validation.rs:431-482
    n = 24
    scene = Scene.empty(Extent.cube(n), 1e-3).with_boundary(Boundary.PEC)
    frequency = scene.grid.frequency_for_resolution(12.0)
    scene.sources.append(Source.point([0.0] * 3, Axis.Z, Waveform.ricker(frequency)))
    sim = Simulation(context, scene)
    sim.advance_by(160)
    ez = sim.component(sim.read_electric(), Axis.Z).reshape(n, n, n)
    sim.close()
    peak = float(np.abs(ez).max())
    assert peak > 0.0
    assert np.all(ez[:, :, 0] == 0.0)
    mirrored = ez[:, :, 1:][:, :, ::-1]
    assert float(np.abs(ez[:, :, 1:] - mirrored).max()) < 1e-6 * peak


def device_probe_trace(context, length: int, steps: int) -> np.ndarray:
    transverse, source_x, probe_x = 40, 20, 36
    scene = Scene.empty(Extent(length, transverse, transverse), 1e-3)
    frequency = scene.grid.frequency_for_resolution(20.0)
    position = [oracle.position(scene.grid, 0, source_x + 0.5),
                oracle.position(scene.grid, 1, 0.5 * transverse),
                oracle.position(scene.grid, 2, 0.5 * transverse)]
    scene.sources.append(Source.point(position, Axis.Z, Waveform.ricker(frequency)))
    sim = Simulation(context, scene)
    index = (transverse // 2 * transverse + transverse // 2) * length + probe_x
    out = []
    for _ in range(steps):
        sim.advance_by(1)
        out.append(sim.component(sim.read_electric(), Axis.Z)[index])
    sim.close()
    return np.array(out, dtype=np.float32)


def test_the_absorbing_layer_reflects_below_the_stated_level_on_the_device(context):
    # tests/validation.rs:523-550: below −50 dB (about −58 measured)
    near, reference = device_probe_trace(context, 72, 380), device_probe_trace(context, 200, 380)
    incident = float(np.abs(reference).max())
    reflected = float(np.abs(near - reference).max())
    decibels = 20.0 * math.log10(reflected / incident)
    print(f"absorbing layer reflection on the device: {decibels:.1f} dB (bar -50 dB)", flush=True)
    assert decibels < -50.0


# ---------------------------------------------------------------------------
# Graded grids (tests/graded.rs), on the device
# ---------------------------------------------------------------------------

def device_trace(context, grid, steps) -> np.ndarray:
    scene = Scene.on_grid(grid)
    frequency = scene.grid.frequency_for_resolution(20.0)
    scene.sources.append(Source.point([-0.012, 0.0, 0.0], Axis.Z, Waveform.ricker(frequency)))
    px, py, pz = oracle.cell_containing(scene.grid, [-0.004, 0.0, 0.0])
    sim = Simulation(context, scene)
    nx, ny, _ = sim.extent
    index = (pz * ny + py) * nx + px
    out = []
    for _ in range(steps):
        sim.advance_by(1)
        out.append(sim.component(sim.read_electric(), Axis.Z)[index])
    sim.close()
    return np.array(out, dtype=np.float32)


def test_stepping_back_survives_a_graded_grid_on_the_device(context):
    # tests/graded.rs:48-93
    scene = Scene.cavity_on(Grid.graded([0.040, 0.028, 0.028], 1.0 / CELL, [band(0.004, FINE)]))
    sim = Simulation(context, scene)
    sim.advance_by(200)
    marked = sim.read_electric()
    excited = sim.energy().total()
    sim.advance_by(150)
    sim.reverse_by(150)
    peak = float(np.abs(marked).max())
    assert peak > 0.0
    assert float(np.abs(marked - sim.read_electric()).max()) < 1e-4 * peak
    sim.reverse_by(200)
    assert sim.energy().total() < 1e-8 * excited
    sim.close()


def test_a_graded_grid_gives_the_same_answer_as_a_uniformly_fine_one_on_the_device(context):
    # tests/graded.rs: the probe trace on a graded grid within 6% of the
    # uniformly fine grid's, resampled to the graded time step
    steps = 300
    graded = Grid.graded([DOMAIN, TRANSVERSE, TRANSVERSE], 1.0 / CELL, [band(0.02, FINE)])
    fine = uniformly_fine()
    gs, fs = graded.time_step(), fine.time_step()
    measured = device_trace(context, graded, steps)
    expected = resample(device_trace(context, fine, int(np.float32(steps) * np.float32(gs) / np.float32(fs)) + 2),
                        fs, len(measured), gs)
    peak = float(np.abs(expected).max())
    assert peak > 0.0
    assert float(np.abs(measured - expected).max()) < 0.06 * peak


def test_a_gentler_grading_transition_reflects_less_on_the_device(context):
    # tests/graded.rs: the default cap of 1.15 reflects below −30 dB, a cap of
    # 1.05 at least 3 dB less
    reference = device_trace(context, uniformly_fine(), 340)
    incident = float(np.abs(reference).max())
    assert incident > 0.0

    def reflection(max_ratio):
        graded = Grid.graded_at([DOMAIN, TRANSVERSE, TRANSVERSE], 1.0 / CELL, [band(0.014, FINE)], max_ratio)
        measured = device_trace(context, graded, 340)
        return 20.0 * math.log10(float(np.abs(measured - reference).max()) / incident)

    default, gentle = reflection(1.15), reflection(1.05)
    print(f"grading transition on the device: {default:.1f} dB at a 1.15 cap, {gentle:.1f} dB at 1.05", flush=True)
    assert gentle < default - 3.0
    assert default < -30.0
