"""The two bandwidth probes: each leaves the sim's state unchanged and
reports a plausible rate, and the stencil kernel cannot outrun its own
tiling run as a copy by more than noise."""
from __future__ import annotations

import numpy as np
import pytest

from diaphane_amd import Extent, Scene, _lib
from diaphane_amd.gpu import Simulation

pytestmark = pytest.mark.gpu


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("rows", [4, 6, 8])
def test_probes_leave_the_state_unchanged(context, rows):
    scene = Scene.photon(Extent(150, 70, 66))  # ragged in x, y and z
    sim = Simulation(context, scene)
    sim.advance_by(30)
    e, h = sim.read_electric(), sim.read_magnetic()
    assert sim.stream_probe(2) > 0.0
    assert sim.tile_probe(rows, 16, 2) > 0.0
    assert bits_equal(sim.read_electric(), e) and bits_equal(sim.read_magnetic(), h)
    sim.advance_by(10)
    again = Simulation(context, scene)
    again.advance_by(40)
    assert bits_equal(sim.read_electric(), again.read_electric())
    with pytest.raises(_lib.DiaphaneError, match="rows"):
        sim.tile_probe(5, 16, 1)
