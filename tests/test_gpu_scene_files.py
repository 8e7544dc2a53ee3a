"""The curated scenes (the reference's scenes/*.ron) on the device against the
oracle, bit for bit: each at full size for 40 steps, and at 400 cells per
metre (the resolution of tests/scenes.rs:113-132) for 300 steps, by which time
every source has fired."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle
from diaphane_amd import Scene
from diaphane_amd.gpu import Simulation

pytestmark = pytest.mark.gpu

SCENES = sorted((Path(__file__).resolve().parent / "golden" / "scenes").glob("*.ron"))


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("resolution,steps", [(None, 40), (400.0, 300)], ids=["full", "coarse"])
@pytest.mark.parametrize("path", SCENES, ids=[p.stem for p in SCENES])
def test_a_curated_scene_matches_the_oracle(context, path, resolution, steps):
    scene = Scene.load(path)
    if resolution:
        scene = scene.with_resolution(resolution)
    print(f"{path.name} {scene.grid.extent} {steps} steps", flush=True)
    reference = oracle.Simulation(scene)
    device = Simulation(context, scene)
    reference.advance_by(steps)
    device.advance_by(steps)
    assert reference.is_finite()
    for name, ref, dev in (("E", reference.electric_all, device.read_electric()),
                           ("H", reference.magnetic_all, device.read_magnetic())):
        diff = np.flatnonzero(ref.view(np.uint32) != dev.view(np.uint32))
        assert diff.size == 0, f"{path.name} {name}: {diff.size} cells differ after {steps} steps"
    assert np.abs(reference.electric_all).max() > 0.0
    device.close()
