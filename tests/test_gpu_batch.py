"""The batched form: members of a dp_batch against the same sims advanced
alone and against the oracle.

Each member of a batch runs the same per-cell arithmetic as a solo sim, so
the bar is bit equality with both (and the 1e-5-of-peak bar, reported).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import AXES, Axis, Boundary, Extent, Scene, Source, Waveform
from diaphane_amd import _lib
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import Simulation
from diaphane_amd.workloads import batched_uneven

from conftest import assert_field_parity

pytestmark = pytest.mark.gpu

TOLERANCE = 1e-5


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def seeded(context, instance):
    sim = Simulation(context, instance.scene)
    e, h = instance.initial_fields()
    sim.write_fields(e, h, 0)
    return sim


def oracle_seeded(instance):
    ref = oracle.Simulation(instance.scene)
    e, h = instance.initial_fields()
    ref.restore_fields(e, h, 0)
    return ref


def assert_matches_oracle(name, reference, candidate, steps):
    assert reference.step_count() == candidate.step_count()
    for label, ref_all, cand in (("E", reference.electric_all, candidate.read_electric()),
                                 ("H", reference.magnetic_all, candidate.read_magnetic())):
        refs = [ref_all[a * reference.cells:(a + 1) * reference.cells] for a in range(3)]
        cands = [candidate.component(cand, a) for a in AXES]
        assert_field_parity(f"{name} {label}", refs, cands, steps, TOLERANCE)
        diff = np.flatnonzero(ref_all.view(np.uint32) != cand.view(np.uint32))
        assert diff.size == 0, f"{name} {label}: {diff.size} cells differ from the oracle after {steps} steps"


def mixed_scenes():
    """Uneven cavity instances plus scenes that inject every step."""
    instances = batched_uneven(5, seed=11, lo=22, hi=70)
    photon = Scene.photon(Extent(70, 30, 26))
    plane = Scene.on_grid(Scene.photon(Extent(40, 24, 24)).grid).with_boundary(Boundary.DEFAULT)
    plane.sources.append(Source.plane_wave(Axis.X, -0.01, Axis.Z, Waveform.gaussian_pulse(
        plane.grid.frequency_for_resolution(16.0), 2.0)))
    cavity = Scene.cavity(Extent(33, 17, 65))
    return instances, [photon, plane, cavity]


@pytest.mark.parametrize("graph_steps,flags", [(16, 2), (0, 2), (16, 18)])
def test_batched_members_match_solo_runs_and_the_oracle(context, graph_steps, flags):
    instances, scenes = mixed_scenes()
    steps = 40  # two graphs of 16 and a tail of 8 when graph_steps = 16
    batch = Batch(context)
    batch.set_graph_steps(graph_steps)
    batch.set_tile(0, 0, flags)
    members = [seeded(context, i) for i in instances] + [Simulation(context, s) for s in scenes]
    for sim in members:
        batch.add(sim)
    batch.advance_by(steps)
    batch.wait()
    assert batch.launch_count() == steps
    solos = [seeded(context, i) for i in instances] + [Simulation(context, s) for s in scenes]
    for solo in solos:
        solo.advance_by(steps)
    for k, (member, solo) in enumerate(zip(members, solos)):
        assert member.step_count() == steps
        assert bits_equal(member.read_electric(), solo.read_electric()), f"member {k}: E differs from solo"
        assert bits_equal(member.read_magnetic(), solo.read_magnetic()), f"member {k}: H differs from solo"
    for k, instance in enumerate(instances[:3]):
        ref = oracle_seeded(instance)
        ref.advance_by(steps)
        assert_matches_oracle(f"instance {k}", ref, members[k], steps)
    for k, scene in enumerate(scenes):
        ref = oracle.Simulation(scene)
        ref.advance_by(steps)
        assert_matches_oracle(scene_name(k), ref, members[len(instances) + k], steps)


def scene_name(k):
    return ("photon", "plane wave", "cavity")[k]


def test_targets_stop_members_at_their_own_step_counts(context):
    instances = batched_uneven(4, seed=12, lo=22, hi=48)
    batch = Batch(context)
    batch.set_graph_steps(8)
    targets = [5, 17, 30, 64]
    members = [seeded(context, i) for i in instances]
    for sim, target in zip(members, targets):
        batch.add(sim, target)
    batch.advance_by(100)
    batch.wait()
    assert batch.launch_count() <= 100
    for sim, target, instance in zip(members, targets, instances):
        assert sim.step_count() == target
        solo = seeded(context, instance)
        solo.advance_by(target)
        assert bits_equal(sim.read_electric(), solo.read_electric())
        assert bits_equal(sim.read_magnetic(), solo.read_magnetic())


def test_members_join_and_leave_between_advances(context):
    instances = batched_uneven(3, seed=13, lo=24, hi=56)
    batch = Batch(context)
    batch.set_graph_steps(8)
    a, b, c = (seeded(context, i) for i in instances)
    batch.add(a)
    batch.add(b)
    batch.advance_by(20)
    batch.remove(b)
    batch.add(c)
    batch.advance_by(12)
    b.advance_by(7)  # a removed member is an ordinary sim again
    batch.wait()
    assert (a.step_count(), b.step_count(), c.step_count()) == (32, 27, 12)
    for sim, instance in ((a, instances[0]), (b, instances[1]), (c, instances[2])):
        solo = seeded(context, instance)
        solo.advance_by(sim.step_count())
        assert bits_equal(sim.read_electric(), solo.read_electric())
        assert bits_equal(sim.read_magnetic(), solo.read_magnetic())


def test_intensity_accumulates_inside_a_batch(context):
    scene = Scene.photon(Extent.cube(32))
    batch = Batch(context)
    member = Simulation(context, scene)
    other = Simulation(context, Scene.cavity(Extent(24, 40, 20)))
    batch.add(member)
    batch.add(other)
    batch.advance_by(60)
    member.set_accumulate_intensity(True)
    batch.advance_by(90)
    batch.wait()
    assert member.accumulated_steps() == 90
    ref = oracle.Simulation(scene)
    ref.advance_by(60)
    ref.set_accumulate_intensity(True)
    ref.advance_by(90)
    assert np.array_equal(ref.intensity_f64(), member.read_intensity_f64())
    assert not np.any(other.read_intensity_f64())


def test_a_member_refuses_solo_advance_and_leaves_when_destroyed(context):
    batch = Batch(context)
    a = Simulation(context, Scene.cavity(Extent.cube(24)))
    b = Simulation(context, Scene.cavity(Extent.cube(20)))
    batch.add(a)
    batch.add(b)
    with pytest.raises(_lib.DiaphaneError, match="batch"):
        a.advance_by(1)
    with pytest.raises(_lib.DiaphaneError, match="already belongs"):
        batch.add(a)
    batch.remove(a)
    a.close()
    b.close()  # destroying a member detaches it
    assert len(batch) == 0
    batch.advance_by(3)
    batch.wait()


@pytest.mark.parametrize("rows,depth,flags", [(6, 16, 2), (8, 5, 0), (4, 64, 2), (4, 32, 18), (6, 7, 16), (4, 16, 22),
                                             (6, 9, 6), (8, 32, 20)])
def test_tile_shapes_do_not_change_results(context, rows, depth, flags):
    # flags 16: one wave per tile instead of the persistent grid; 4: buffer addressing
    instances = batched_uneven(3, seed=14, lo=22, hi=80)
    batch = Batch(context)
    batch.set_tile(rows, depth, flags)
    members = [seeded(context, i) for i in instances]
    for sim in members:
        batch.add(sim)
    batch.advance_by(25)
    batch.wait()
    for sim, instance in zip(members, instances):
        solo = seeded(context, instance)
        solo.advance_by(25)
        assert bits_equal(sim.read_electric(), solo.read_electric())
        assert bits_equal(sim.read_magnetic(), solo.read_magnetic())


def test_autotuning_leaves_members_unchanged(context):
    instances = batched_uneven(4, seed=15, lo=22, hi=90)
    batch = Batch(context)
    members = [seeded(context, i) for i in instances]
    for sim in members:
        batch.add(sim)
    batch.advance_by(9)
    rows, depth, ms = batch.autotune(2)
    # 10: two steps per launch; depths are powers of two or three quarters / one and a half of one
    assert rows in (4, 6, 8, 10) and depth in (6, 8, 12, 16, 24, 32, 48, 64, 96) and ms > 0.0
    batch.advance_by(11)
    batch.wait()
    for sim, instance in zip(members, instances):
        assert sim.step_count() == 20
        solo = seeded(context, instance)
        solo.advance_by(20)
        assert bits_equal(sim.read_electric(), solo.read_electric())
        assert bits_equal(sim.read_magnetic(), solo.read_magnetic())
