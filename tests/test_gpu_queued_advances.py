"""Advances queued without waiting, against the oracle bit for bit.

An `advance` call outside the sim's injection window uploads the
descriptors of its next 64 steps (or more) into one of a few event-guarded
slots and returns without waiting for the previous call; calls inside the
window reuse them. Every step must still see its own descriptors: the fields
after many back-to-back one- and two-step calls (windows crossed, tables of
different sizes, a time-dependent source injecting on every step, a restore
back into a window, other uploads coming round the slots) equal the oracle's
exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Material, Scene, Shape, Source, Waveform
from diaphane_amd.gpu import TILE_BUFFER, TILE_NONTEMPORAL, TILE_TWO_STEP, Simulation

pytestmark = pytest.mark.gpu


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("rows,depth,flags", [(4, 0, TILE_NONTEMPORAL), (8, 0, TILE_TWO_STEP | TILE_BUFFER)])
def test_back_to_back_short_advances_match_the_oracle(context, rows, depth, flags):
    scene = Scene.photon(Extent(48, 32, 40))
    reference = oracle.Simulation(scene)
    device = Simulation(context, scene)
    device.set_tile(rows, depth, flags)
    chunks = [1] * 23 + [2, 3, 1, 2] * 4 + [1] * 9  # 64 calls, 72 steps
    for chunk in chunks:
        device.advance_by(chunk)
    reference.advance_by(sum(chunks))
    assert device.step_count() == reference.step_count() == sum(chunks)
    assert np.abs(reference.electric_all).max() > 0.0
    assert bits_equal(reference.electric_all, device.read_electric())
    assert bits_equal(reference.magnetic_all, device.read_magnetic())
    device.close()


def test_a_long_table_between_short_ones_and_a_reversal(context):
    # A reversible scene: queued forward calls (one long table among short
    # ones), then queued reverse calls, which take their descriptors by value.
    scene = Scene.empty(Extent(40, 36, 36), 1e-3).with_boundary(Boundary.PEC)
    frequency = scene.grid.frequency_for_resolution(24.0)
    glass = scene.materials.push(Material.refractive(2.0))
    scene.shapes.append(Shape.Sphere([4e-3, 0.0, 0.0], 6e-3, glass))
    scene.sources.append(Source.point([-12e-3, 0.0, 0.0], Axis.Z, Waveform.ricker(frequency)))
    scene.validate()
    reference = oracle.Simulation(scene)
    device = Simulation(context, scene)
    forward = [1, 1, 1, 97, 1, 1, 5, 1, 1, 1, 1]
    for chunk in forward:
        device.advance_by(chunk)
    reference.advance_by(sum(forward))
    assert bits_equal(reference.electric_all, device.read_electric())
    for chunk in [1, 1, 3, 1, 40]:
        device.reverse_by(chunk)
        device.advance_by(1)
        reference.reverse_by(chunk)
        reference.advance_by(1)
    assert device.step_count() == reference.step_count()
    assert bits_equal(reference.electric_all, device.read_electric())
    assert bits_equal(reference.magnetic_all, device.read_magnetic())
    device.close()


def test_the_injection_window_survives_restores_and_other_uploads(context):
    # A sim keeps the uploaded descriptors of 64 steps and reuses them across
    # advances. A restore back into the window, autotuner uploads that come
    # round to the window's slot of the ring, and advances crossing the
    # window's end must all still apply each step's own descriptors.
    scene = Scene.photon(Extent(48, 32, 40))
    reference = oracle.Simulation(scene)
    device = Simulation(context, scene)
    tile = (8, 0, TILE_TWO_STEP | TILE_BUFFER)
    device.set_tile(*tile)
    device.advance_by(10)
    e10, h10 = device.read_electric(), device.read_magnetic()
    device.advance_by(20)
    device.write_fields(e10, h10, 10)  # back to step 10, inside the window of steps 1-64
    for _ in range(5):
        device.autotune(1)  # five more uploads into a ring of four slots
    device.set_tile(*tile)
    for chunk in [3, 1, 50, 7, 1]:  # steps 11-72, across the window's end
        device.advance_by(chunk)
    reference.advance_by(72)
    assert device.step_count() == 72
    assert np.abs(reference.electric_all).max() > 0.0
    assert bits_equal(reference.electric_all, device.read_electric())
    assert bits_equal(reference.magnetic_all, device.read_magnetic())
    device.close()
