"""The batched and served workloads themselves, at their stated sizes.

batched_uneven_256: all 256 uneven solves (sides log-uniform in [48, 160] per
axis, seeded cavity modes) advanced together on the default batch for an odd
number of steps; every member is bit-equal to the same solve run alone, and
eight sampled members to the oracle.

served_stream: requests drawn with the stream's stated side range [32, 256]
and its three scene kinds (step counts cut to 200..400 so the test stays
short) go through the server while it is busy with earlier ones; every result
is bit-equal to the request run alone, and a sample to the oracle.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import oracle
from diaphane_amd import workloads
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import Simulation
from diaphane_amd.serve import DONE, Server

pytestmark = pytest.mark.gpu


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def seeded(context, instance):
    sim = Simulation(context, instance.scene)
    e, h = instance.initial_fields()
    sim.write_fields(e, h, 0)
    return sim


def test_the_256_uneven_solves_batched_match_their_solo_runs(context):
    instances = workloads.batched_uneven_256()
    assert len(instances) == 256
    steps = 23  # odd: the default two-step batch ends every member on a single step
    batch = Batch(context)
    members = [seeded(context, i) for i in instances]
    for sim in members:
        batch.add(sim)
    print(f"batch of {len(batch)}: {batch.cells()} cells, tiling {batch.tiling()}", flush=True)
    batch.advance_by(steps)
    batch.wait()
    sample = set(range(0, 256, 32))
    for k, (instance, member) in enumerate(zip(instances, members)):
        assert member.step_count() == steps
        solo = seeded(context, instance)
        solo.advance_by(steps)
        e, h = member.read_electric(), member.read_magnetic()
        assert bits_equal(e, solo.read_electric()), f"member {k}: E differs from its solo run"
        assert bits_equal(h, solo.read_magnetic()), f"member {k}: H differs from its solo run"
        solo.close()
        if k in sample:
            reference = oracle.Simulation(instance.scene)
            e0, h0 = instance.initial_fields()
            reference.restore_fields(e0, h0, 0)
            reference.advance_by(steps)
            assert bits_equal(e, reference.electric_all), f"member {k}: E differs from the oracle"
            assert bits_equal(h, reference.magnetic_all), f"member {k}: H differs from the oracle"
    batch.close()
    for sim in members:
        sim.close()


def test_the_served_stream_at_its_stated_sizes_matches_solo_runs(context):
    stream = workloads.served_stream(36, rate=50.0, seed=5, lo=32, hi=256, steps_lo=200, steps_hi=400)
    sides = [max(r.scene.grid.extent.as_array()) for r in stream]
    assert max(sides) > 160, "the draw should reach the upper end of the side range"
    assert {r.kind for r in stream} == set(workloads.SERVED_KINDS)
    with Server(context, chunk_steps=32, host_threads=4) as server:
        tickets = []
        start = time.perf_counter()
        for request in stream:
            # keep the arrival gaps, so requests join a batch already stepping
            delay = request.arrival_s - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            tickets.append(server.submit(request.scene, request.steps))
        for ticket in tickets:
            server.wait(ticket)
        joined_busy = server.stats()["chunks"]
        print(f"{len(stream)} requests served in {joined_busy} chunks", flush=True)
        sampled = 0
        for k, (request, ticket) in enumerate(zip(stream, tickets)):
            info = server.info(ticket)
            assert info.state == DONE and info.status == 0 and info.step == request.steps
            result = server.result(ticket)
            e, h = result.read_electric(), result.read_magnetic()
            solo = Simulation(context, request.scene)
            solo.advance_by(request.steps)
            assert bits_equal(e, solo.read_electric()), f"request {k} ({request.kind}): E differs from alone"
            assert bits_equal(h, solo.read_magnetic()), f"request {k} ({request.kind}): H differs from alone"
            solo.close()
            # the oracle for one request of each kind, the smaller ones
            if sampled < 3 and request.scene.grid.extent.total() < 2_000_000 and k % 3 == sampled:
                reference = oracle.Simulation(request.scene)
                reference.advance_by(request.steps)
                assert bits_equal(e, reference.electric_all), f"request {k}: E differs from the oracle"
                assert bits_equal(h, reference.magnetic_all), f"request {k}: H differs from the oracle"
                sampled += 1
            server.release(ticket)
        assert sampled == 3
