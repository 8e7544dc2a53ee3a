"""The unit tests of src/timeline.rs (:217-421), ported: the host keyframe
ring (diaphane_amd.timeline.Timeline) over the oracle's Steppable. The device
ring is held to the same tests in test_gpu_timeline.py."""
from __future__ import annotations

import numpy as np

import oracle
from diaphane_amd import Boundary, Extent, Scene
from diaphane_amd.timeline import Seek, Timeline


def scene() -> Scene:
    return Scene.photon(Extent.cube(24))


def fingerprint(sim) -> np.ndarray:
    return sim.electric_all.copy()


def test_a_snapshot_restores_the_exact_state():
    sim = oracle.Simulation(scene())
    sim.advance_by(60)
    snapshot = sim.snapshot()
    expected = fingerprint(sim)
    sim.advance_by(40)
    assert not np.array_equal(fingerprint(sim), expected)
    sim.restore(snapshot)
    assert sim.step_count() == 60
    assert np.array_equal(fingerprint(sim).view(np.uint32), expected.view(np.uint32))


def test_seeking_backwards_reproduces_the_forward_run_exactly():
    reference = oracle.Simulation(scene())
    reference.advance_by(90)
    sim = oracle.Simulation(scene())
    timeline = Timeline(25, 8)
    for _ in range(150):
        sim.advance_by(1)
        timeline.observe(sim)
    outcome = timeline.seek(sim, 90)
    assert sim.step_count() == 90
    assert np.array_equal(fingerprint(sim).view(np.uint32), fingerprint(reference).view(np.uint32))
    assert outcome == Seek("from_keyframe", 15, 75)


def test_a_target_older_than_the_window_replays_from_the_start():
    sim = oracle.Simulation(scene())
    timeline = Timeline(20, 2)
    for _ in range(120):
        sim.advance_by(1)
        timeline.observe(sim)
    assert timeline.keyframe_count() == 2
    assert timeline.earliest() == 100
    reference = oracle.Simulation(scene())
    reference.advance_by(10)
    assert timeline.seek(sim, 10) == Seek("from_start", 10)
    assert np.array_equal(fingerprint(sim), fingerprint(reference))


def test_seeking_forward_just_steps():
    sim = oracle.Simulation(scene())
    timeline = Timeline(10, 4)
    sim.advance_by(20)
    assert timeline.seek(sim, 35) == Seek("forward", 15)
    assert sim.step_count() == 35
    assert timeline.seek(sim, 35) == Seek("unchanged")


def test_the_window_slides_and_the_memory_bill_is_bounded():
    sim = oracle.Simulation(scene())
    timeline = Timeline(10, 4)
    for _ in range(200):
        sim.advance_by(1)
        timeline.observe(sim)
    assert timeline.keyframe_count() == 4
    assert timeline.earliest() == 170
    assert timeline.bytes() == 4 * 24 * Extent.cube(24).total()


def test_observing_twice_at_the_same_step_records_one_keyframe():
    sim = oracle.Simulation(scene())
    timeline = Timeline(10, 8)
    sim.advance_by(10)
    timeline.observe(sim)
    timeline.observe(sim)
    assert timeline.keyframe_count() == 1


def test_batches_that_never_land_on_a_multiple_still_record():
    sim = oracle.Simulation(scene())
    timeline = Timeline(10, 8)
    sim.advance_by(7)
    for _ in range(20):
        sim.advance_by(3)
        timeline.observe(sim)
    assert timeline.keyframe_count() >= 5, f"only {timeline.keyframe_count()} keyframes over 67 steps"


def test_recording_survives_a_seek_to_an_unaligned_step():
    sim = oracle.Simulation(scene())
    timeline = Timeline(10, 8)
    for _ in range(6):
        sim.advance_by(8)
        timeline.observe(sim)
    timeline.seek(sim, 13)
    before = timeline.keyframe_count()
    for _ in range(10):
        sim.advance_by(8)
        timeline.observe(sim)
    assert timeline.keyframe_count() > before, "recording stopped after seeking to an unaligned step"
    assert timeline.seek(sim, 90).kind == "from_keyframe"


def test_a_snapshot_reports_its_own_size():
    sim = oracle.Simulation(scene())
    assert sim.snapshot().bytes() == 24 * Extent.cube(24).total()


def test_scrubbing_a_lossless_scene_agrees_with_stepping_backwards():
    lossless = Scene.cavity(Extent.cube(24)).with_boundary(Boundary.PEC)
    scrubbed = oracle.Simulation(lossless)
    timeline = Timeline(20, 8)
    for _ in range(100):
        scrubbed.advance_by(1)
        timeline.observe(scrubbed)
    reversed_ = oracle.Simulation(lossless)
    reversed_.advance_by(100)
    reversed_.reverse_by(30)
    timeline.seek(scrubbed, 70)
    assert scrubbed.step_count() == reversed_.step_count()
    peak = float(np.abs(fingerprint(scrubbed)).max())
    worst = float(np.abs(fingerprint(scrubbed) - fingerprint(reversed_)).max())
    assert worst < 1e-4 * peak, f"replay and reversal disagree by {worst / peak:e} of the peak"
