"""The viewer's auto-range reduction (dp_sim_measure_peak) against the same
maxima taken in numpy over the oracle's fields (examples/common/render.rs:
382-419, shaders/volume.wgsl:473-493). Maxima of f32 values computed in the
same order are exact, so the bar is equality."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Extent, Scene
from diaphane_amd.gpu import Simulation

pytestmark = pytest.mark.gpu


def expected_peaks(electric, magnetic, cells, intensity=None, steps=0):
    e = [electric[a * cells:(a + 1) * cells] for a in range(3)]
    h = [magnetic[a * cells:(a + 1) * cells] for a in range(3)]
    half = np.float32(0.5)
    ue = half * (((np.float32(0) + e[0] * e[0]) + e[1] * e[1]) + e[2] * e[2])
    uh = half * (((np.float32(0) + h[0] * h[0]) + h[1] * h[1]) + h[2] * h[2])
    peaks = [max(ue.max(), uh.max())] + [np.abs(c).max() for c in e] + [np.abs(c).max() for c in h]
    if intensity is not None and steps:
        peaks.append((intensity.astype(np.float32) / np.float32(steps)).max())
    else:
        peaks.append(np.float32(0))
    return np.array(peaks, dtype=np.float32)


@pytest.mark.parametrize("scene", [Scene.photon(Extent(48, 32, 32)), Scene.cavity(Extent(40, 36, 32))],
                         ids=["photon", "cavity"])
def test_the_peaks_match_the_oracle(context, scene):
    device, reference = Simulation(context, scene), oracle.Simulation(scene)
    for sim in (device, reference):
        sim.advance_by(60)
    assert np.array_equal(device.measure_peak(), expected_peaks(reference.electric_all, reference.magnetic_all,
                                                                reference.cells))
    for sim in (device, reference):
        sim.set_accumulate_intensity(True)
        sim.advance_by(30)
    got = device.measure_peak()
    want = expected_peaks(reference.electric_all, reference.magnetic_all, reference.cells,
                          reference.intensity_f64(), 30)
    assert np.array_equal(got, want), (got, want)
    assert got[7] > 0.0


def test_a_large_grid_reduces_across_many_blocks(context):
    scene = Scene.photon(Extent(256, 200, 180))
    device = Simulation(context, scene)
    device.advance_by(40)
    got = device.measure_peak()
    want = expected_peaks(device.read_electric(), device.read_magnetic(), device.cells)
    assert np.array_equal(got, want)
    assert got[0] > 0.0
