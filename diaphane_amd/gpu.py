"""The device solver: gpu::Simulation's counterpart (src/gpu.rs:159-799).

A thin binding over the C ABI; every call goes to libdiaphane_amd.so, which
runs the hand-written gfx950 kernels. There is no CPU fallback: without a
device, ``headless_context`` raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_uint32, c_void_p
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ._lib import DiaphaneError, check, lib
from .scene import Axis, Scene
from .timeline import Snapshot

KERNEL_FUSED = 0
KERNEL_TWO_PASS = 1
KERNEL_FUSED_OVERLAP = 2
TILE_XCD_REMAP = 1
TILE_NONTEMPORAL = 2
TILE_BUFFER = 4
TILE_TWO_STEP = 8
TILE_STACK = 32  # with TILE_TWO_STEP: waves stacked along y; waves (2 or 4) in bits 16..19

BUFFER_ELECTRIC = 0
BUFFER_MAGNETIC = 1
BUFFER_INTENSITY = 2
BUFFER_MATERIAL_INDEX = 3
BUFFER_INTENSITY_F32 = 4
BUFFER_LOOKUP = 5


@dataclass
class Energy:
    """cpu::Energy (src/cpu.rs:52-62)."""

    electric: float
    magnetic: float

    def total(self) -> float:
        return self.electric + self.magnetic


class Context:
    """An opened HIP device (headless_context, src/gpu.rs:805-817)."""

    def __init__(self, device: int = 0):
        handle = c_void_p()
        check(lib().dp_engine_create(device, byref(handle)))
        self._handle = handle

    @property
    def handle(self) -> c_void_p:
        return self._handle

    def arch(self) -> str:
        buf = ctypes.create_string_buffer(64)
        check(lib().dp_engine_arch(self._handle, buf, 64))
        return buf.value.decode()

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle and lib is not None:
            try:
                lib().dp_engine_destroy(handle)
            except Exception:
                pass
            self._handle = None


_default_context: Optional[Context] = None


def headless_context(device: int = 0) -> Context:
    """Opens (once) and returns the context for ``device``."""
    global _default_context
    if _default_context is None:
        _default_context = Context(device)
    return _default_context


class PinnedBuffer:
    """Page-locked host memory from the engine (dp_host_alloc), freed when
    this object goes; `array` views it."""

    def __init__(self, count: int, dtype=np.float32):
        dtype = np.dtype(dtype)
        ptr = c_void_p()
        check(lib().dp_host_alloc(int(count) * dtype.itemsize, byref(ptr)))
        self._ptr = ptr
        raw = (ctypes.c_char * (int(count) * dtype.itemsize)).from_address(ptr.value)
        self.array = np.frombuffer(raw, dtype=dtype, count=int(count))

    def __del__(self):
        ptr = getattr(self, "_ptr", None)
        if ptr:
            try:
                lib().dp_host_free(ptr)
            except Exception:
                pass
            self._ptr = None


def _fptr(array: np.ndarray):
    return array.ctypes.data_as(POINTER(c_float))


class Simulation:
    """A running simulation on the device."""

    TIMEOUT_MS = 120_000

    _read_targets: tuple = ()  # host buffers of the readbacks in flight

    def __init__(self, context: Context, scene: Scene):
        self._context = context
        self._scene_c = scene.to_c()
        handle = c_void_p()
        check(lib().dp_sim_create(context.handle, byref(self._scene_c), byref(handle)))
        self._handle = handle
        self.grid = scene.grid
        extent = (c_uint32 * 3)()
        check(lib().dp_sim_extent(self._handle, extent))
        self.extent = (int(extent[0]), int(extent[1]), int(extent[2]))
        self.cells = self.extent[0] * self.extent[1] * self.extent[2]

    @classmethod
    def borrowed(cls, handle: c_void_p, grid=None) -> "Simulation":
        """A view of a sim owned elsewhere (a served request's result); closing
        it leaves the sim alone."""
        self = cls.__new__(cls)
        self._context = None
        self._scene_c = None
        self._handle = handle
        self._owned = False
        self.grid = grid
        extent = (c_uint32 * 3)()
        check(lib().dp_sim_extent(self._handle, extent))
        self.extent = (int(extent[0]), int(extent[1]), int(extent[2]))
        self.cells = self.extent[0] * self.extent[1] * self.extent[2]
        return self

    def close(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle and getattr(self, "_owned", True):
            lib().dp_sim_destroy(handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference surface ------------------------------------------------
    def step_count(self) -> int:
        return int(lib().dp_sim_step_count(self._handle))

    def time(self) -> float:
        return float(lib().dp_sim_time(self._handle))

    def time_step(self) -> float:
        return float(lib().dp_sim_time_step(self._handle))

    def field_bytes(self) -> int:
        return int(lib().dp_sim_field_bytes(self._handle))

    def bytes_per_step(self) -> int:
        return int(lib().dp_sim_bytes_per_step(self._handle))

    def is_reversible(self) -> bool:
        return bool(lib().dp_sim_is_reversible(self._handle))

    def advance_by(self, steps: int) -> None:
        check(lib().dp_sim_advance(self._handle, int(steps)))

    def reverse_by(self, steps: int) -> None:
        check(lib().dp_sim_reverse(self._handle, int(steps)))

    def wait(self, timeout_ms: int = 0) -> None:
        check(lib().dp_sim_wait(self._handle, int(timeout_ms)))

    def reset(self) -> None:
        check(lib().dp_sim_reset(self._handle))

    def set_accumulate_intensity(self, on: bool) -> None:
        check(lib().dp_sim_set_accumulate_intensity(self._handle, 1 if on else 0))

    def accumulated_steps(self) -> int:
        return int(lib().dp_sim_accumulated_steps(self._handle))

    def _field_out(self, out: Optional[np.ndarray]) -> np.ndarray:
        if out is None:
            return np.empty(3 * self.cells, dtype=np.float32)
        if out.dtype != np.float32 or out.size != 3 * self.cells or not out.flags.c_contiguous:
            raise DiaphaneError(-4, "out must be a contiguous float32 array of 3 x cell count")
        return out

    def read_electric(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        """E, component-major; into `out` when given (a reused host buffer
        skips the page faults of a fresh one)."""
        out = self._field_out(out)
        check(lib().dp_sim_read_electric(self._handle, _fptr(out), out.size))
        return out

    def read_magnetic(self, out: Optional[np.ndarray] = None) -> np.ndarray:
        out = self._field_out(out)
        check(lib().dp_sim_read_magnetic(self._handle, _fptr(out), out.size))
        return out

    def begin_read(self, which: int, out: np.ndarray) -> None:
        """Starts copying E (BUFFER_ELECTRIC) or H (BUFFER_MAGNETIC) as of
        this call into `out` and returns at once; later advances run while
        the copy does. Two readbacks can be in flight (with two in flight
        this first waits for the older). `out` (a PinnedBuffer's array for
        the copy to overlap) must stay untouched until its readback has
        landed: wait_reads(1) after the next begin_read, or finish_read."""
        out = self._field_out(out)
        check(lib().dp_sim_begin_read(self._handle, int(which), _fptr(out), out.size))
        # kept alive while the copies may write them
        self._read_targets = (self._read_targets + (out,))[-2:]

    def wait_reads(self, in_flight: int, timeout_ms: int = 0) -> None:
        """Waits until at most `in_flight` readbacks are in flight, the
        oldest landing first."""
        check(lib().dp_sim_wait_reads(self._handle, int(in_flight), int(timeout_ms)))
        keep = min(int(in_flight), len(self._read_targets))
        self._read_targets = self._read_targets[len(self._read_targets) - keep:]

    def finish_read(self, timeout_ms: int = 0) -> None:
        """Waits for every readback in flight."""
        check(lib().dp_sim_finish_read(self._handle, int(timeout_ms)))
        self._read_targets = ()

    def read_intensity(self) -> np.ndarray:
        out = np.empty(self.cells, dtype=np.float32)
        check(lib().dp_sim_read_intensity(self._handle, _fptr(out), out.size))
        return out

    def read_intensity_f64(self) -> np.ndarray:
        out = np.empty(self.cells, dtype=np.float64)
        check(lib().dp_sim_read_intensity_f64(self._handle, out.ctypes.data_as(POINTER(c_double)), out.size))
        return out

    def component(self, values: np.ndarray, axis: Axis) -> np.ndarray:
        a = int(axis)
        return values[a * self.cells:(a + 1) * self.cells]

    def write_fields(self, electric: np.ndarray, magnetic: np.ndarray, step: int) -> None:
        e = np.ascontiguousarray(electric, dtype=np.float32)
        h = np.ascontiguousarray(magnetic, dtype=np.float32)
        if e.size != h.size:
            raise DiaphaneError(-4, "snapshot was taken from a different grid")
        check(lib().dp_sim_write_fields(self._handle, _fptr(e), _fptr(h), e.size, int(step)))

    def energy(self) -> Energy:
        ue, uh = c_double(), c_double()
        check(lib().dp_sim_energy(self._handle, byref(ue), byref(uh)))
        return Energy(ue.value, uh.value)

    def device_ptr(self, which: int) -> int:
        out = c_void_p()
        check(lib().dp_sim_device_ptr(self._handle, which, byref(out)))
        return int(out.value or 0)

    def lookup_samples(self) -> int:
        """lookup_samples (src/gpu.rs:349-351): samples per axis in the
        renderer's world-to-cell table (device_ptr(BUFFER_LOOKUP))."""
        return int(lib().dp_sim_lookup_samples(self._handle))

    # -- Steppable (src/timeline.rs:63-72) ---------------------------------
    def snapshot(self) -> Snapshot:
        return Snapshot(self.step_count(), self.read_electric(), self.read_magnetic())

    def restore(self, snapshot: Snapshot) -> None:
        self.write_fields(snapshot.electric, snapshot.magnetic, snapshot.step)

    # -- engine controls ---------------------------------------------------
    def set_kernel(self, kernel: int) -> None:
        check(lib().dp_sim_set_kernel(self._handle, kernel))

    def set_tile(self, rows: int = 0, depth: int = 0, flags: int = -1) -> None:
        """Fused-kernel tiling; flags: 1 = XCD-contiguous tiles, 2 = non-temporal
        stores, 4 = buffer addressing, 8 = two steps per launch."""
        check(lib().dp_sim_set_tile(self._handle, rows, depth, flags))

    def autotune(self, trials: int = 3) -> dict:
        """Times the fused kernel's tile candidates on this grid and keeps the
        fastest; the fields are unchanged."""
        rows, depth, flags, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), c_float()
        check(lib().dp_sim_autotune(self._handle, trials, byref(rows), byref(depth), byref(flags), byref(ms)))
        return {"rows": rows.value, "depth": depth.value, "flags": flags.value, "ms_per_step": ms.value}

    def record_event(self, slot: int) -> None:
        check(lib().dp_sim_record_event(self._handle, slot))

    def elapsed_ms(self, start: int, end: int) -> float:
        ms = c_float()
        check(lib().dp_sim_elapsed_ms(self._handle, start, end, byref(ms)))
        return float(ms.value)

    def stream_probe(self, iterations: int = 10) -> float:
        """GB/s of the step's access mix streamed with no stencil."""
        gbs = c_double()
        check(lib().dp_sim_stream_probe(self._handle, iterations, byref(gbs)))
        return gbs.value

    def measure_peak(self) -> np.ndarray:
        """The viewer's auto-range maxima (8 floats): max(½|E|², ½|H|²), max
        |E_x..z|, max |H_x..z|, and the peak averaged intensity."""
        out = np.zeros(8, dtype=np.float32)
        check(lib().dp_sim_measure_peak(self._handle, _fptr(out)))
        return out

    def tile_probe(self, rows: int = 0, depth: int = 0, iterations: int = 10) -> float:
        """GB/s (49 B/cell) of the fused kernel's tile access pattern with no
        stencil, at `rows` × `depth` (0: the plain tiling's)."""
        gbs = c_double()
        check(lib().dp_sim_tile_probe(self._handle, rows, depth, iterations, byref(gbs)))
        return gbs.value

    def launch_count(self) -> int:
        return int(lib().dp_sim_launch_count(self._handle))
