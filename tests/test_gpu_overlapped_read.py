"""Overlapped readbacks (dp_sim_begin_read / dp_sim_wait_reads /
dp_sim_finish_read).

begin_read captures E or H as of the call, after every step queued before
it, and copies it to the host while later advances run. The host buffer
must then hold exactly the oracle's fields at that step, bit for bit, for
one and two steps per launch (where the launches after the call overwrite
the buffer the fields were read from), with two readbacks in flight and a
third waiting for the older, for a batch member while the batch keeps
stepping, and when the sim is destroyed with the copy in flight.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Extent, Scene
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import (BUFFER_ELECTRIC, BUFFER_INTENSITY, BUFFER_MAGNETIC, TILE_BUFFER, TILE_STACK,
                              TILE_TWO_STEP, DiaphaneError, PinnedBuffer, Simulation)

pytestmark = pytest.mark.gpu


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def oracle_at(scene: Scene, steps: int):
    reference = oracle.Simulation(scene)
    reference.advance_by(steps)
    return reference.electric_all.copy(), reference.magnetic_all.copy()


@pytest.mark.parametrize("tile", [None, (8, 2, TILE_TWO_STEP), (4, 2, TILE_TWO_STEP | TILE_STACK | TILE_BUFFER)],
                         ids=["default", "two-step", "stacked"])
def test_readbacks_hold_the_fields_of_their_step(context, tile):
    scene = Scene.photon(Extent(70, 36, 30))
    sim = Simulation(context, scene)
    if tile:
        sim.set_tile(*tile)
    e_buf, h_buf = PinnedBuffer(3 * sim.cells), PinnedBuffer(3 * sim.cells)
    sim.advance_by(37)
    sim.begin_read(BUFFER_ELECTRIC, e_buf.array)
    sim.advance_by(9)  # several launches write both buffers of the pair meanwhile
    sim.finish_read()
    e37, _ = oracle_at(scene, 37)
    assert float(np.abs(e37).max()) > 0.0
    assert same_bits(e_buf.array, e37)
    sim.begin_read(BUFFER_MAGNETIC, h_buf.array)
    sim.finish_read()
    e46, h46 = oracle_at(scene, 46)
    assert same_bits(h_buf.array, h46)
    # a second readback begun while one is in flight runs behind it
    sim.begin_read(BUFFER_ELECTRIC, e_buf.array)
    sim.advance_by(3)
    sim.begin_read(BUFFER_MAGNETIC, h_buf.array)
    sim.finish_read()
    assert same_bits(e_buf.array, e46)
    _, h49 = oracle_at(scene, 49)
    assert same_bits(h_buf.array, h49)
    sim.finish_read()  # nothing in flight: returns at once
    sim.close()


@pytest.mark.parametrize("tile", [None, (8, 2, TILE_TWO_STEP), (4, 2, TILE_TWO_STEP | TILE_STACK | TILE_BUFFER)],
                         ids=["default", "two-step", "stacked"])
def test_two_readbacks_in_flight_land_in_order(context, tile):
    scene = Scene.photon(Extent(70, 36, 30))
    sim = Simulation(context, scene)
    if tile:
        sim.set_tile(*tile)
    bufs = [PinnedBuffer(3 * sim.cells) for _ in range(3)]
    sim.advance_by(11)
    sim.begin_read(BUFFER_ELECTRIC, bufs[0].array)
    sim.advance_by(6)
    sim.begin_read(BUFFER_MAGNETIC, bufs[1].array)
    sim.wait_reads(1)  # the older of the two has landed
    e11, _ = oracle_at(scene, 11)
    assert float(np.abs(e11).max()) > 0.0
    assert same_bits(bufs[0].array, e11)
    sim.advance_by(3)
    sim.begin_read(BUFFER_ELECTRIC, bufs[2].array)  # takes the slot the first one freed
    sim.advance_by(5)
    sim.finish_read()
    _, h17 = oracle_at(scene, 17)
    e20, _ = oracle_at(scene, 20)
    assert same_bits(bufs[1].array, h17)
    assert same_bits(bufs[2].array, e20)
    sim.wait_reads(5)  # nothing in flight: returns at once
    sim.close()


def test_a_readback_begun_with_two_in_flight_waits_for_the_older(context):
    scene = Scene.photon(Extent(52, 40, 34))
    sim = Simulation(context, scene)
    bufs = [PinnedBuffer(3 * sim.cells) for _ in range(5)]
    done = 0
    for k, buf in enumerate(bufs):
        sim.advance_by(4 + k)
        done += 4 + k
        sim.begin_read(BUFFER_ELECTRIC, buf.array)  # from the third on, waits for the oldest
    sim.advance_by(3)
    sim.finish_read()
    done = 0
    for k, buf in enumerate(bufs):
        done += 4 + k
        expected, _ = oracle_at(scene, done)
        assert same_bits(buf.array, expected), f"readback {k} at step {done}"
    sim.close()


def test_a_batch_member_reads_back_while_the_batch_steps(context):
    scenes = [Scene.photon(Extent(48, 40, 28)), Scene.cavity(Extent(36, 30, 44))]
    sims = [Simulation(context, s) for s in scenes]
    batch = Batch(context)
    for sim in sims:
        batch.add(sim)
    batch.advance_by(21)
    buf = PinnedBuffer(3 * sims[0].cells)
    sims[0].begin_read(BUFFER_ELECTRIC, buf.array)
    batch.advance_by(14)
    sims[0].finish_read()
    batch.wait()
    e21, _ = oracle_at(scenes[0], 21)
    assert same_bits(buf.array, e21)
    e35, _ = oracle_at(scenes[0], 35)
    assert same_bits(sims[0].read_electric(), e35)
    batch.close()
    for sim in sims:
        sim.close()


def test_destroying_the_sim_waits_for_the_copy(context):
    scene = Scene.photon(Extent(64, 48, 40))
    sim = Simulation(context, scene)
    sim.advance_by(25)
    buf = PinnedBuffer(3 * sim.cells)
    sim.begin_read(BUFFER_ELECTRIC, buf.array)
    sim.advance_by(6)
    sim.close()
    e25, _ = oracle_at(scene, 25)
    assert same_bits(buf.array, e25)


def test_bad_readback_requests_are_refused(context):
    sim = Simulation(context, Scene.photon(Extent(24, 24, 24)))
    out = np.empty(3 * sim.cells, dtype=np.float32)
    with pytest.raises(DiaphaneError):
        sim.begin_read(BUFFER_INTENSITY, out)
    with pytest.raises(DiaphaneError):
        sim.begin_read(BUFFER_ELECTRIC, np.empty(3 * sim.cells - 1, dtype=np.float32))
    sim.finish_read()
    sim.close()
