"""The oracle's intensity accumulator (src/cpu.rs:186-196, 577-598) against
its closed form: after each complete step (E-phase injections included) every
cell gains s = ((0 + Ex²) + Ey²) + Ez² formed in f32, summed in f32 as the
reference keeps it and in f64 as the engine keeps it. Both are recomputed
here with numpy from the oracle's E after each step and must agree bit for
bit; switching on starts a fresh sum, and restore, reverse and reset clear
it."""
from __future__ import annotations

import numpy as np

import oracle
from diaphane_amd import Axis, Boundary, Extent, Grid, Scene, Source, Waveform


def step_sum(sim) -> np.ndarray:
    e = sim.electric_all
    n = sim.cells
    ex, ey, ez = e[:n], e[n:2 * n], e[2 * n:]
    s = np.float32(0.0) + ex * ex
    s = s + ey * ey
    return s + ez * ez


class Track:
    def __init__(self, cells: int):
        self.f32 = np.zeros(cells, dtype=np.float32)
        self.f64 = np.zeros(cells, dtype=np.float64)
        self.steps = 0

    def add(self, s: np.ndarray) -> None:
        self.f32 = self.f32 + s
        self.f64 = self.f64 + s.astype(np.float64)
        self.steps += 1


def assert_track(sim, track: Track) -> None:
    assert sim.accumulated_steps() == track.steps
    assert np.array_equal(sim.intensity().view(np.uint32), track.f32.view(np.uint32))
    assert np.array_equal(sim.intensity_f64().view(np.uint64), track.f64.view(np.uint64))


def point_scene(boundary=Boundary.DEFAULT) -> Scene:
    grid = Grid.new(Extent(30, 26, 22), 1e-3)
    scene = Scene(grid, boundary)
    frequency = grid.frequency_for_resolution(12.0)
    # A soft point source: its E-phase injection lands on a cell after the E
    # sweep, so a sum taken before the injections would miss it there.
    scene.sources.append(Source.point([1.5e-3, -2.5e-3, 0.5e-3], Axis.Z, Waveform.ricker(frequency)))
    scene.sources.append(Source.point([-4.5e-3, 3.5e-3, -1.5e-3], Axis.X, Waveform.ricker(1.3 * frequency)))
    return scene


def test_each_step_adds_the_square_of_the_complete_field():
    sim = oracle.Simulation(point_scene())
    sim.advance_by(5)
    sim.set_accumulate_intensity(True)
    track = Track(sim.cells)
    for _ in range(14):
        sim.advance()
        track.add(step_sum(sim))
        assert_track(sim, track)
    assert track.f64.max() > 0.0
    # the f32 sum rounds where the f64 one does not
    assert np.any(track.f32.astype(np.float64) != track.f64)


def test_the_source_cell_sees_its_own_injection():
    scene = point_scene()
    sim = oracle.Simulation(scene)
    sim.advance_by(9)
    sim.set_accumulate_intensity(True)
    sim.advance()
    source_cell = oracle.cell_containing(scene.grid, scene.sources[0].at)
    nx, ny, _ = sim.extent
    index = (source_cell[2] * ny + source_cell[1]) * nx + source_cell[0]
    assert sim.intensity_f64()[index] == float(step_sum(sim)[index]) > 0.0


def test_switching_off_holds_and_switching_on_again_starts_fresh():
    sim = oracle.Simulation(point_scene())
    sim.set_accumulate_intensity(True)
    track = Track(sim.cells)
    for _ in range(6):
        sim.advance()
        track.add(step_sum(sim))
    sim.set_accumulate_intensity(False)
    sim.advance_by(4)
    assert_track(sim, track)  # off: nothing added, nothing cleared
    sim.set_accumulate_intensity(True)  # on again: a fresh average
    fresh = Track(sim.cells)
    assert_track(sim, fresh)
    for _ in range(3):
        sim.advance()
        fresh.add(step_sum(sim))
    sim.set_accumulate_intensity(True)  # already on: keeps its sum
    sim.advance()
    fresh.add(step_sum(sim))
    assert_track(sim, fresh)


def test_restore_reverse_and_reset_clear_the_sum():
    scene = point_scene(Boundary.PEC)
    sim = oracle.Simulation(scene)
    sim.advance_by(3)
    snapshot = sim.snapshot()
    sim.set_accumulate_intensity(True)
    sim.advance_by(5)
    assert sim.accumulated_steps() == 5
    sim.restore(snapshot)
    assert_track(sim, Track(sim.cells))
    sim.advance_by(4)
    sim.reverse_by(2)
    assert_track(sim, Track(sim.cells))
    sim.advance_by(2)
    assert sim.accumulated_steps() == 2
    sim.reset()
    assert_track(sim, Track(sim.cells))
