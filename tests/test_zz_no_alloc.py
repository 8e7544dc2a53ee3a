"""dp_sim_measure_peak_noalloc, the entry point that allocates nothing,
against dp_sim_measure_peak bit for bit: two calls in a row, the second on a
smaller grid, through the C ABI."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from diaphane_amd import Extent, Scene, _lib
from diaphane_amd.gpu import Simulation

pytestmark = pytest.mark.gpu


def peaks(sim: Simulation):
    lib = _lib.lib()
    allocating = np.zeros(8, dtype=np.float32)
    fixed = np.zeros(8, dtype=np.float32)
    pointer = ctypes.POINTER(ctypes.c_float)
    _lib.check(lib.dp_sim_measure_peak(sim._handle, allocating.ctypes.data_as(pointer)))
    _lib.check(lib.dp_sim_measure_peak_noalloc(sim._handle, fixed.ctypes.data_as(pointer)))
    return allocating, fixed


def test_the_fixed_scratch_gives_the_allocating_result(context):
    # A large grid first (many partial blocks), then a smaller one: the second
    # call must not see the first call's partials.
    large = Simulation(context, Scene.slab(Extent(256, 200, 180), 1.5))
    large.advance_by(40)
    large.set_accumulate_intensity(True)
    large.advance_by(9)
    small = Simulation(context, Scene.photon(Extent(40, 32, 30)))
    small.advance_by(25)
    for label, sim in (("256x200x180 slab", large), ("40x32x30 photon", small)):
        allocating, fixed = peaks(sim)
        print(f"{label}: {fixed}", flush=True)
        assert np.all(allocating[:7] > 0.0), f"{label}: the fields are still zero"
        assert np.array_equal(allocating.view(np.uint32), fixed.view(np.uint32)), \
            f"{label}: {fixed} vs {allocating}"
    assert peaks(large)[1][7] > 0.0 and peaks(small)[1][7] == 0.0  # only the large one accumulates
    large.close()
    small.close()


def test_repeated_calls_on_one_sim_agree(context):
    sim = Simulation(context, Scene.cavity(Extent(48, 44, 40)))
    sim.advance_by(30)
    first = peaks(sim)[1]
    for _ in range(3):
        allocating, fixed = peaks(sim)
        assert np.array_equal(fixed.view(np.uint32), first.view(np.uint32))
        assert np.array_equal(allocating.view(np.uint32), fixed.view(np.uint32))
    sim.close()
