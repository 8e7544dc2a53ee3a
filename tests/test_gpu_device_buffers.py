"""Device buffers for a renderer (dp_sim_device_ptr): what a viewer binding
the engine's memory in the same process reads, against the readbacks."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from diaphane_amd import Extent, Scene
from diaphane_amd.gpu import (BUFFER_ELECTRIC, BUFFER_INTENSITY, BUFFER_INTENSITY_F32, BUFFER_MAGNETIC,
                              BUFFER_MATERIAL_INDEX, Simulation)

pytestmark = pytest.mark.gpu


def device_read(ptr: int, count: int, dtype) -> np.ndarray:
    # The HIP runtime the engine already loaded in this process.
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(count, dtype=dtype)
    status = hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                           ctypes.c_size_t(out.nbytes), ctypes.c_int(2))  # device to host
    assert status == 0, f"hipMemcpy failed: {status}"
    return out


def test_device_buffers_hold_what_the_readbacks_return(context):
    sim = Simulation(context, Scene.slab(Extent(40, 32, 30), 1.5))
    sim.advance_by(7)
    sim.set_accumulate_intensity(True)
    sim.advance_by(11)
    sim.wait()
    n = sim.cells
    e = device_read(sim.device_ptr(BUFFER_ELECTRIC), 3 * n, np.float32)
    h = device_read(sim.device_ptr(BUFFER_MAGNETIC), 3 * n, np.float32)
    assert np.array_equal(e.view(np.uint32), sim.read_electric().view(np.uint32))
    assert np.array_equal(h.view(np.uint32), sim.read_magnetic().view(np.uint32))
    wide = device_read(sim.device_ptr(BUFFER_INTENSITY), n, np.float64)
    assert np.array_equal(wide, sim.read_intensity_f64()) and wide.max() > 0.0
    ptr = sim.device_ptr(BUFFER_INTENSITY_F32)
    sim.wait()
    narrow = device_read(ptr, n, np.float32)
    assert np.array_equal(narrow.view(np.uint32), sim.read_intensity().view(np.uint32))
    assert np.array_equal(narrow, wide.astype(np.float32))
    material = device_read(sim.device_ptr(BUFFER_MATERIAL_INDEX), n, np.uint8)
    assert set(np.unique(material).tolist()) == {0, 1}
    # the f32 view is a snapshot: refreshed by the next request, not by stepping
    sim.advance_by(3)
    sim.wait()
    assert np.array_equal(device_read(ptr, n, np.float32).view(np.uint32), narrow.view(np.uint32))
    sim.device_ptr(BUFFER_INTENSITY_F32)
    sim.wait()
    assert np.array_equal(device_read(ptr, n, np.float32).view(np.uint32), sim.read_intensity().view(np.uint32))
    sim.close()


def test_the_f32_view_of_a_sim_that_never_accumulated_is_zero(context):
    sim = Simulation(context, Scene.cavity(Extent(24, 20, 22)))
    sim.advance_by(5)
    ptr = sim.device_ptr(BUFFER_INTENSITY_F32)
    sim.wait()
    assert not np.any(device_read(ptr, sim.cells, np.float32))
    sim.close()
