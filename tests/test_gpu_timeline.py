"""The keyframe ring in device memory (dp_timeline) — the tests of
src/timeline.rs:231-422, run on the device and held to the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Extent, Scene, _lib
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import Simulation
from diaphane_amd.timeline import DeviceTimeline, Seek

pytestmark = pytest.mark.gpu


def scene():
    return Scene.photon(Extent.cube(24))


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def oracle_after(steps: int):
    reference = oracle.Simulation(scene())
    reference.advance_by(steps)
    return reference


def test_seeking_backwards_reproduces_the_forward_run_exactly(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(25, 8)
    for _ in range(150):
        simulation.advance_by(1)
        timeline.observe(simulation)
    outcome = timeline.seek(simulation, 90)
    assert simulation.step_count() == 90
    assert outcome == Seek("from_keyframe", 15, 75)
    reference = oracle_after(90)
    assert bits_equal(simulation.read_electric(), reference.electric_all)
    assert bits_equal(simulation.read_magnetic(), reference.magnetic_all)


def test_a_target_older_than_the_window_replays_from_the_start(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(20, 2)
    for _ in range(120):
        simulation.advance_by(1)
        timeline.observe(simulation)
    assert timeline.keyframe_count() == 2
    assert timeline.earliest() == 100
    assert timeline.seek(simulation, 10) == Seek("from_start", 10)
    assert bits_equal(simulation.read_electric(), oracle_after(10).electric_all)


def test_seeking_forward_just_steps(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(10, 4)
    simulation.advance_by(20)
    assert timeline.seek(simulation, 35) == Seek("forward", 15)
    assert simulation.step_count() == 35
    assert timeline.seek(simulation, 35) == Seek("unchanged")


def test_the_window_slides_and_the_memory_bill_is_bounded(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(10, 4)
    for _ in range(200):
        simulation.advance_by(1)
        timeline.observe(simulation)
    assert timeline.keyframe_count() == 4
    assert timeline.earliest() == 170
    assert timeline.bytes() == 4 * 24 * Extent.cube(24).total()


def test_observing_twice_at_the_same_step_records_one_keyframe(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(10, 8)
    simulation.advance_by(10)
    timeline.observe(simulation)
    timeline.observe(simulation)
    assert timeline.keyframe_count() == 1


def test_batches_that_never_land_on_a_multiple_still_record(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(10, 8)
    simulation.advance_by(7)
    for _ in range(20):
        simulation.advance_by(3)
        timeline.observe(simulation)
    assert timeline.keyframe_count() >= 5


def test_recording_survives_a_seek_to_an_unaligned_step(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(10, 8)
    for _ in range(6):
        simulation.advance_by(8)
        timeline.observe(simulation)
    timeline.seek(simulation, 13)
    before = timeline.keyframe_count()
    for _ in range(10):
        simulation.advance_by(8)
        timeline.observe(simulation)
    assert timeline.keyframe_count() > before
    outcome = timeline.seek(simulation, 90)
    assert outcome.kind == "from_keyframe"
    assert bits_equal(simulation.read_electric(), oracle_after(90).electric_all)


def test_a_restore_clears_the_intensity_sum_and_scrubbing_matches_the_host_ring(context):
    # The device ring restores what the host ring (snapshot/write_fields) does.
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(10, 16)
    simulation.set_accumulate_intensity(True)
    for _ in range(12):
        simulation.advance_by(5)
        timeline.observe(simulation)
    assert simulation.accumulated_steps() == 60
    timeline.seek(simulation, 37)
    assert simulation.accumulated_steps() == 7  # restored at 30, replayed 7
    electric = simulation.read_electric()
    host = Simulation(context, scene())
    host.advance_by(30)
    frame = host.snapshot()
    host.advance_by(20)
    host.set_accumulate_intensity(True)
    host.restore(frame)
    host.advance_by(7)
    assert bits_equal(electric, host.read_electric())
    assert np.array_equal(simulation.read_intensity_f64(), host.read_intensity_f64())


def test_one_ring_serves_two_sims_of_the_same_grid(context):
    first = Simulation(context, scene())
    timeline = DeviceTimeline(16, 4)
    for _ in range(4):
        first.advance_by(16)
        timeline.observe(first)
    second = Simulation(context, scene())
    second.advance_by(70)
    assert timeline.seek(second, 50) == Seek("from_keyframe", 2, 48)
    assert bits_equal(second.read_electric(), oracle_after(50).electric_all)


def test_the_ring_refuses_another_grid_and_batch_members(context):
    simulation = Simulation(context, scene())
    timeline = DeviceTimeline(5, 2)
    simulation.advance_by(10)
    timeline.observe(simulation)
    other = Simulation(context, Scene.photon(Extent.cube(28)))
    other.advance_by(20)
    with pytest.raises(_lib.DiaphaneError, match="different grid"):
        timeline.observe(other)
    batch = Batch(context)
    member = Simulation(context, scene())
    batch.add(member)
    with pytest.raises(_lib.DiaphaneError, match="batch"):
        timeline.observe(member)
    batch.close()


def test_a_snapshot_restores_the_exact_state(context):
    # src/timeline.rs:232-245, through Steppable::snapshot/restore on the device
    simulation = Simulation(context, scene())
    simulation.advance_by(60)
    snapshot = simulation.snapshot()
    expected = simulation.read_electric()
    simulation.advance_by(40)
    assert not bits_equal(simulation.read_electric(), expected)
    simulation.restore(snapshot)
    assert simulation.step_count() == 60
    assert bits_equal(simulation.read_electric(), expected)
    assert bits_equal(expected, oracle_after(60).electric_all)
    assert snapshot.bytes() == 24 * Extent.cube(24).total()  # src/timeline.rs:385-390
    simulation.close()


def test_scrubbing_a_lossless_scene_agrees_with_stepping_backwards(context):
    # src/timeline.rs:392-420: replay from a device keyframe and the exact
    # inverse update land on the same state; each route is also the oracle's.
    from diaphane_amd import Boundary

    lossless = Scene.cavity(Extent.cube(24)).with_boundary(Boundary.PEC)
    scrubbed = Simulation(context, lossless)
    timeline = DeviceTimeline(20, 8)
    for _ in range(100):
        scrubbed.advance_by(1)
        timeline.observe(scrubbed)
    reversed_ = Simulation(context, lossless)
    reversed_.advance_by(100)
    reversed_.reverse_by(30)
    timeline.seek(scrubbed, 70)
    assert scrubbed.step_count() == reversed_.step_count() == 70
    a, b = scrubbed.read_electric(), reversed_.read_electric()
    peak = float(np.abs(a).max())
    assert float(np.abs(a - b).max()) < 1e-4 * peak
    forward = oracle.Simulation(lossless)
    forward.advance_by(70)
    assert bits_equal(a, forward.electric_all)
    backward = oracle.Simulation(lossless)
    backward.advance_by(100)
    backward.reverse_by(30)
    assert bits_equal(b, backward.electric_all)
    for sim in (scrubbed, reversed_):
        sim.close()
    timeline.close()
