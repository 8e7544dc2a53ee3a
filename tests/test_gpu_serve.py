"""The served form: each request's result is the same alone and in the mix.

Requests of a stream are run alone through the plain single-sim path and
then through a server, staggered so that they join a batch that is already
stepping others; every result must match its solo run bit for bit (and the
oracle, for a sample).
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import oracle
from diaphane_amd import AXES, Boundary, Extent, Grid, Scene
from diaphane_amd import _lib
from diaphane_amd.gpu import Simulation
from diaphane_amd.serve import DONE, Server
from diaphane_amd.workloads import served_stream

from conftest import assert_field_parity

pytestmark = pytest.mark.gpu


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def small_stream(count=9, seed=21):
    return served_stream(count, rate=200.0, seed=seed, lo=32, hi=56, steps_lo=60, steps_hi=160)


def solo_fields(context, request):
    sim = Simulation(context, request.scene)
    sim.advance_by(request.steps)
    fields = (sim.read_electric(), sim.read_magnetic())
    sim.close()
    return fields


@pytest.mark.parametrize("one_at_a_time", [False, True])
def test_each_result_is_the_same_alone_and_in_the_mix(context, one_at_a_time):
    stream = small_stream()
    alone = [solo_fields(context, r) for r in stream]
    with Server(context, chunk_steps=16, host_threads=3, one_at_a_time=one_at_a_time) as server:
        tickets = []
        for k, request in enumerate(stream):
            tickets.append(server.submit(request.scene, request.steps))
            if k % 3 == 2:
                time.sleep(0.02)  # later requests join a batch already stepping
        for ticket in tickets:
            server.wait(ticket)
        for k, (request, ticket) in enumerate(zip(stream, tickets)):
            info = server.info(ticket)
            assert info.state == DONE and info.status == 0
            assert info.step == request.steps
            assert info.submitted_s <= info.admitted_s <= info.started_s <= info.completed_s
            result = server.result(ticket)
            assert result.step_count() == request.steps
            e, h = alone[k]
            assert bits_equal(result.read_electric(), e), f"request {k} ({request.kind}): E differs from alone"
            assert bits_equal(result.read_magnetic(), h), f"request {k} ({request.kind}): H differs from alone"
            server.release(ticket)
        assert server.stats()["used_bytes"] == 0


def test_results_read_while_other_requests_run(context):
    # A client copying a finished result to the host while the loop keeps
    # capturing and launching graphs for the others: the library's copies
    # are stream-ordered, never on the legacy stream (which HIP refuses
    # while a stream is capturing).
    stream = served_stream(8, rate=200.0, seed=25, lo=32, hi=48, steps_lo=40, steps_hi=1200)
    alone = [solo_fields(context, r) for r in stream]
    with Server(context, chunk_steps=8) as server:
        tickets = {server.submit(r.scene, r.steps): k for k, r in enumerate(stream)}
        read_while_running = 0
        while tickets:
            for ticket in list(tickets):
                if server.info(ticket).state == DONE:
                    k = tickets.pop(ticket)
                    result = server.result(ticket)
                    read_while_running += server.stats()["running"] > 0
                    assert bits_equal(result.read_electric(), alone[k][0]), f"request {k}: E differs"
                    assert bits_equal(result.read_magnetic(), alone[k][1]), f"request {k}: H differs"
                    server.release(ticket)
            time.sleep(0.001)
        assert read_while_running > 0


def test_served_results_match_the_oracle(context):
    stream = small_stream(3, seed=22)
    with Server(context, chunk_steps=8) as server:
        tickets = [server.submit(r.scene, r.steps) for r in stream]
        for request, ticket in zip(stream, tickets):
            server.wait(ticket)
            result = server.result(ticket)
            ref = oracle.Simulation(request.scene)
            ref.advance_by(request.steps)
            for label, ref_all, cand in (("E", ref.electric_all, result.read_electric()),
                                         ("H", ref.magnetic_all, result.read_magnetic())):
                refs = [ref_all[a * ref.cells:(a + 1) * ref.cells] for a in range(3)]
                cands = [result.component(cand, a) for a in AXES]
                if max(float(np.abs(r).max()) for r in refs) > 2e-3:
                    assert_field_parity(f"{request.kind} {label}", refs, cands, request.steps, 1e-5)
                assert bits_equal(ref_all, cand), f"{request.kind} {label} differs from the oracle"
            server.release(ticket)


def test_admission_holds_to_the_budget(context):
    stream = small_stream(6, seed=23)
    biggest = max(r.scene.grid.extent.total() for r in stream)
    budget = int(biggest * 49 * 1.5) + (2 << 20)  # room for one request at a time, sometimes two
    with Server(context, hbm_budget_bytes=budget, chunk_steps=16) as server:
        tickets = [server.submit(r.scene, r.steps) for r in stream]
        peak = 0
        pending = set(tickets)
        while pending:
            stats = server.stats()
            peak = max(peak, stats["used_bytes"])
            assert stats["used_bytes"] <= budget
            for ticket in list(pending):
                if server.info(ticket).state == DONE:
                    server.wait(ticket)
                    server.release(ticket)
                    pending.discard(ticket)
            time.sleep(0.002)
        assert peak > 0
    with Server(context, hbm_budget_bytes=1 << 20) as server:
        with pytest.raises(_lib.DiaphaneError, match="budget"):
            server.submit(stream[0].scene, 10)


def test_a_failing_request_reports_its_error(context):
    # an absorbing layer that does not fit the grid: the sim cannot be built
    scene = Scene(Grid(Extent(16, 16, 16), 1e-3), Boundary.DEFAULT)
    with Server(context) as server:
        bad = server.submit(scene, 10)
        good_request = small_stream(1, seed=24)[0]
        good = server.submit(good_request.scene, good_request.steps)
        with pytest.raises(_lib.DiaphaneError, match="absorbing"):
            server.wait(bad)
        server.wait(good)
        server.release(bad)
        server.release(good)
        assert server.stats()["used_bytes"] == 0


@pytest.mark.parametrize("chunk_cells", [1, Server.ALL_MEMBERS])
def test_the_chunk_schedule_does_not_change_results(context, chunk_cells):
    # chunk_cells = 1: one member (the least remaining work, or one that has
    # waited too long) steps per chunk; ALL_MEMBERS: every member steps.
    stream = small_stream(6, seed=26)
    alone = [solo_fields(context, r) for r in stream]
    with Server(context, chunk_steps=8, chunk_cells=chunk_cells) as server:
        tickets = [server.submit(r.scene, r.steps) for r in stream]
        for ticket in tickets:
            server.wait(ticket)
        completed = []
        for k, (request, ticket) in enumerate(zip(stream, tickets)):
            info = server.info(ticket)
            assert info.state == DONE and info.step == request.steps
            completed.append((info.completed_s, request.scene.grid.extent.total() * request.steps))
            result = server.result(ticket)
            assert bits_equal(result.read_electric(), alone[k][0]), f"request {k}: E differs"
            assert bits_equal(result.read_magnetic(), alone[k][1]), f"request {k}: H differs"
            server.release(ticket)
    if chunk_cells == 1:
        # least work first: the smallest request finishes no later than the largest
        smallest = min(completed, key=lambda c: c[1])
        largest = max(completed, key=lambda c: c[1])
        assert smallest[0] <= largest[0]
