"""Batches advanced two steps per launch (DP_TILE_TWO_STEP on a dp_batch).

Each launch advances every member two steps, or one on the last step of an
odd count. The members must stay bit-identical to the same sims advanced
alone and to the oracle: across graph chunks with odd tails, per-member
targets, intensity switched on at an odd step, and members joining and
leaving between advances.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Extent, Scene
from diaphane_amd import _lib
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import TILE_BUFFER, TILE_NONTEMPORAL, TILE_STACK, TILE_TWO_STEP, TILE_XCD_REMAP, Simulation
from diaphane_amd.workloads import batched_uneven

from test_gpu_batch import assert_matches_oracle, bits_equal, mixed_scenes, oracle_seeded, scene_name, seeded

pytestmark = pytest.mark.gpu

SHAPES = [(8, 0, TILE_TWO_STEP | TILE_NONTEMPORAL), (6, 16, TILE_TWO_STEP), (10, 9, TILE_TWO_STEP | TILE_BUFFER),
          # XCD-grouped workgroup order: runs of 8, and of 3 (a tail past the last whole group)
          (6, 0, TILE_TWO_STEP | TILE_BUFFER | TILE_XCD_REMAP), (8, 5, TILE_TWO_STEP | TILE_XCD_REMAP | (3 << 8)),
          # waves stacked along y
          (4, 0, TILE_TWO_STEP | TILE_BUFFER | TILE_STACK | TILE_XCD_REMAP | (4 << 8)),
          (6, 7, TILE_TWO_STEP | TILE_STACK | (2 << 16))]


@pytest.mark.parametrize("graph_steps", [16, 0])
@pytest.mark.parametrize("rows,depth,flags", SHAPES)
def test_two_step_batches_match_solo_runs_and_the_oracle(context, graph_steps, rows, depth, flags):
    instances, scenes = mixed_scenes()
    steps = 41  # with 16-step graphs: two full chunks and a tail of 9 (four pairs and a single)
    batch = Batch(context)
    batch.set_graph_steps(graph_steps)
    batch.set_tile(rows, depth, flags)
    members = [seeded(context, i) for i in instances] + [Simulation(context, s) for s in scenes]
    for sim in members:
        batch.add(sim)
    batch.advance_by(steps)
    batch.wait()
    assert batch.launch_count() == (8 + 8 + 5 if graph_steps else 21)
    solos = [seeded(context, i) for i in instances] + [Simulation(context, s) for s in scenes]
    for solo in solos:
        solo.advance_by(steps)
    for k, (member, solo) in enumerate(zip(members, solos)):
        assert member.step_count() == steps
        assert bits_equal(member.read_electric(), solo.read_electric()), f"member {k}: E differs from solo"
        assert bits_equal(member.read_magnetic(), solo.read_magnetic()), f"member {k}: H differs from solo"
    for k, instance in enumerate(instances[:2]):
        ref = oracle_seeded(instance)
        ref.advance_by(steps)
        assert_matches_oracle(f"instance {k}", ref, members[k], steps)
    for k, scene in enumerate(scenes):
        ref = oracle.Simulation(scene)
        ref.advance_by(steps)
        assert_matches_oracle(scene_name(k), ref, members[len(instances) + k], steps)


def test_odd_targets_end_with_a_single_step(context):
    instances = batched_uneven(4, seed=12, lo=22, hi=48)
    batch = Batch(context)
    batch.set_graph_steps(8)
    batch.set_tile(8, 0, TILE_TWO_STEP)
    targets = [5, 17, 30, 63]
    members = [seeded(context, i) for i in instances]
    for sim, target in zip(members, targets):
        batch.add(sim, target)
    batch.advance_by(100)
    batch.wait()
    for sim, target, instance in zip(members, targets, instances):
        assert sim.step_count() == target
        solo = seeded(context, instance)
        solo.advance_by(target)
        assert bits_equal(sim.read_electric(), solo.read_electric())
        assert bits_equal(sim.read_magnetic(), solo.read_magnetic())


def test_intensity_switched_on_at_an_odd_step(context):
    scene = Scene.photon(Extent.cube(32))
    batch = Batch(context)
    batch.set_tile(6, 0, TILE_TWO_STEP)
    member = Simulation(context, scene)
    other = Simulation(context, Scene.cavity(Extent(24, 40, 20)))
    batch.add(member)
    batch.add(other)
    batch.advance_by(61)
    member.set_accumulate_intensity(True)
    batch.advance_by(91)
    batch.wait()
    assert member.accumulated_steps() == 91
    ref = oracle.Simulation(scene)
    ref.advance_by(61)
    ref.set_accumulate_intensity(True)
    ref.advance_by(91)
    assert np.array_equal(ref.intensity_f64(), member.read_intensity_f64())
    assert bits_equal(ref.electric_all, member.read_electric())
    assert not np.any(other.read_intensity_f64())


def test_members_join_and_leave_a_two_step_batch(context):
    instances = batched_uneven(3, seed=13, lo=24, hi=56)
    batch = Batch(context)
    batch.set_graph_steps(8)
    batch.set_tile(10, 0, TILE_TWO_STEP | TILE_BUFFER)
    a, b, c = (seeded(context, i) for i in instances)
    batch.add(a)
    batch.add(b)
    batch.advance_by(21)
    batch.remove(b)
    batch.add(c)
    batch.advance_by(13)
    b.advance_by(7)  # a removed member is an ordinary sim again
    batch.wait()
    assert (a.step_count(), b.step_count(), c.step_count()) == (34, 28, 13)
    for sim, instance in ((a, instances[0]), (b, instances[1]), (c, instances[2])):
        solo = seeded(context, instance)
        solo.advance_by(sim.step_count())
        assert bits_equal(sim.read_electric(), solo.read_electric())
        assert bits_equal(sim.read_magnetic(), solo.read_magnetic())


def test_the_batch_autotuner_times_two_step_shapes_and_leaves_members_alone(context):
    instances = batched_uneven(4, seed=15, lo=40, hi=120)
    batch = Batch(context)
    members = [seeded(context, i) for i in instances]
    for sim in members:
        batch.add(sim)
    batch.advance_by(3)
    batch.wait()
    before = [(sim.read_electric(), sim.read_magnetic()) for sim in members]
    batch.autotune(1)
    for sim, (e, h) in zip(members, before):
        assert bits_equal(e, sim.read_electric()) and bits_equal(h, sim.read_magnetic())
    batch.advance_by(9)
    batch.wait()
    for sim, instance in zip(members, instances):
        solo = seeded(context, instance)
        solo.advance_by(12)
        assert bits_equal(sim.read_electric(), solo.read_electric())


def test_two_step_rows_are_checked(context):
    batch = Batch(context)
    with pytest.raises(_lib.DiaphaneError, match="6, 8 or 10"):
        batch.set_tile(4, 0, TILE_TWO_STEP)
    with pytest.raises(_lib.DiaphaneError, match="4, 6 or 8"):
        batch.set_tile(10, 0, 0)
