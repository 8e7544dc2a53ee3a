"""CPU oracle: ctypes binding of liboracle.so, the plain-C restatement of
kvark/diaphane's src/cpu.rs and scene layer (see oracle.h).

TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the benchmark's CPU
baseline use it as the checker. The engine never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, byref, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

from diaphane_amd.scene import DpGridSpec, DpInjection, DpScene, DpWaveform, Scene

HERE = Path(__file__).resolve().parent
LIBRARY_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    """Compiles liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIBRARY_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            build()
        l = ctypes.CDLL(str(LIBRARY_PATH))
        sim = c_void_p
        decl = {
            "or_last_error": (ctypes.c_char_p, []),
            "or_grid_resolve": (c_int, [POINTER(DpGridSpec), POINTER(c_uint32), POINTER(c_float), c_size_t,
                                        POINTER(c_float)]),
            "or_grid_dual": (c_int, [POINTER(DpGridSpec), POINTER(c_float), c_size_t]),
            "or_grid_packed_geometry": (c_int, [POINTER(DpGridSpec), POINTER(c_float), c_size_t]),
            "or_grid_coordinate": (c_float, [POINTER(DpGridSpec), c_int, c_float]),
            "or_grid_position": (c_float, [POINTER(DpGridSpec), c_int, c_float]),
            "or_grid_cell_lookup": (c_int, [POINTER(DpGridSpec), c_uint32, POINTER(c_float), c_size_t]),
            "or_grid_cell_containing": (c_int, [POINTER(DpGridSpec), POINTER(c_float), POINTER(c_uint32)]),
            "or_grid_frequency_for_resolution": (c_float, [POINTER(DpGridSpec), c_float]),
            "or_grid_worst_ratio": (c_float, [POINTER(DpGridSpec)]),
            "or_numerical_phase_velocity": (c_float, [c_float, c_float, POINTER(c_float)]),
            "or_material_coefficients": (c_int, [c_void_p, c_uint32, c_float, POINTER(c_float)]),
            "or_absorber_packed": (c_int, [POINTER(DpScene), POINTER(c_float), c_size_t]),
            "or_waveform_evaluate": (c_float, [POINTER(DpWaveform), c_float]),
            "or_scene_injections": (c_int, [POINTER(DpScene), c_float, POINTER(DpInjection), c_size_t]),
            "or_injection_weight": (c_float, [POINTER(DpScene), POINTER(DpInjection), POINTER(c_uint32)]),
            "or_scene_material_indices": (c_int, [POINTER(DpScene), POINTER(c_uint32), c_size_t]),
            "or_sim_create": (sim, [POINTER(DpScene)]),
            "or_sim_destroy": (None, [sim]),
            "or_sim_set_threads": (None, [sim, c_int]),
            "or_sim_cells": (c_uint64, [sim]),
            "or_sim_extent": (None, [sim, POINTER(c_uint32)]),
            "or_sim_step_count": (c_uint64, [sim]),
            "or_sim_time": (c_float, [sim]),
            "or_sim_advance": (None, [sim, c_uint64]),
            "or_sim_is_reversible": (c_int, [sim]),
            "or_sim_reverse": (c_int, [sim, c_uint64]),
            "or_sim_reset": (None, [sim]),
            "or_sim_set_accumulate_intensity": (None, [sim, c_int]),
            "or_sim_accumulated_steps": (c_uint64, [sim]),
            "or_sim_electric": (POINTER(c_float), [sim]),
            "or_sim_magnetic": (POINTER(c_float), [sim]),
            "or_sim_intensity": (POINTER(c_float), [sim]),
            "or_sim_intensity_f64": (POINTER(c_double), [sim]),
            "or_sim_restore": (c_int, [sim, POINTER(c_float), POINTER(c_float), c_size_t, c_uint64]),
            "or_sim_f64_begin": (c_int, [sim]),
            "or_sim_f64_advance": (c_int, [sim, c_uint64]),
            "or_sim_electric_f64": (POINTER(c_double), [sim]),
            "or_sim_magnetic_f64": (POINTER(c_double), [sim]),
            "or_sim_step_count_f64": (c_uint64, [sim]),
            "or_sim_energy": (None, [sim, POINTER(c_double), POINTER(c_double)]),
            "or_sim_is_finite": (c_int, [sim]),
        }
        for name, (restype, argtypes) in decl.items():
            fn = getattr(l, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = l
    return _lib


class OracleError(RuntimeError):
    pass


def _err() -> str:
    m = lib().or_last_error()
    return m.decode() if m else ""


def _check(status: int) -> None:
    if status != 0:
        raise OracleError(_err())


# -- scene-layer functions --------------------------------------------------
def grid_resolve(grid):
    spec = grid.spec()
    extent = (c_uint32 * 3)()
    _check(lib().or_grid_resolve(byref(spec), extent, None, 0, None))
    n = sum(int(v) for v in extent)
    widths = np.zeros(n, dtype=np.float32)
    dt = c_float()
    _check(lib().or_grid_resolve(byref(spec), extent, widths.ctypes.data_as(POINTER(c_float)), n, byref(dt)))
    return tuple(int(v) for v in extent), widths, float(dt.value)


def grid_dual(grid) -> np.ndarray:
    spec = grid.spec()
    extent, _, _ = grid_resolve(grid)
    out = np.zeros(sum(extent), dtype=np.float32)
    _check(lib().or_grid_dual(byref(spec), out.ctypes.data_as(POINTER(c_float)), out.size))
    return out


def packed_geometry(grid) -> np.ndarray:
    spec = grid.spec()
    extent, _, _ = grid_resolve(grid)
    out = np.zeros(3 * sum(extent), dtype=np.float32)
    _check(lib().or_grid_packed_geometry(byref(spec), out.ctypes.data_as(POINTER(c_float)), out.size))
    return out


def coordinate(grid, axis: int, position: float) -> float:
    return float(lib().or_grid_coordinate(byref(grid.spec()), int(axis), float(position)))


def position(grid, axis: int, coordinate_value: float) -> float:
    return float(lib().or_grid_position(byref(grid.spec()), int(axis), float(coordinate_value)))


def cell_lookup(grid, samples: int) -> np.ndarray:
    """Grid::cell_lookup (src/grid.rs:635-646)."""
    out = np.zeros(3 * int(samples), dtype=np.float32)
    _check(lib().or_grid_cell_lookup(byref(grid.spec()), int(samples), out.ctypes.data_as(POINTER(c_float)),
                                     out.size))
    return out


def cell_containing(grid, at) -> tuple:
    p = (c_float * 3)(*[float(v) for v in at])
    out = (c_uint32 * 3)()
    _check(lib().or_grid_cell_containing(byref(grid.spec()), p, out))
    return tuple(int(v) for v in out)


def frequency_for_resolution(grid, cells: float) -> float:
    return float(lib().or_grid_frequency_for_resolution(byref(grid.spec()), float(cells)))


def worst_ratio(grid) -> float:
    return float(lib().or_grid_worst_ratio(byref(grid.spec())))


def numerical_phase_velocity(courant: float, cells: float, direction) -> float:
    d = (c_float * 3)(*[float(v) for v in direction])
    return float(lib().or_numerical_phase_velocity(float(courant), float(cells), d))


def material_coefficients(scene: Scene) -> np.ndarray:
    c = scene.to_c()
    dt = grid_resolve(scene.grid)[2]
    out = np.zeros(4 * c.material_count, dtype=np.float32)
    _check(lib().or_material_coefficients(ctypes.cast(c.materials, c_void_p), c.material_count, dt,
                                          out.ctypes.data_as(POINTER(c_float))))
    return out.reshape(-1, 4)


def absorber_packed(scene: Scene) -> np.ndarray:
    c = scene.to_c()
    extent, _, _ = grid_resolve(scene.grid)
    out = np.zeros(2 * sum(extent), dtype=np.float32)
    _check(lib().or_absorber_packed(byref(c), out.ctypes.data_as(POINTER(c_float)), out.size))
    return out


def waveform_evaluate(waveform, time: float) -> float:
    w = waveform.to_c()
    return float(lib().or_waveform_evaluate(byref(w), float(time)))


def scene_injections(scene: Scene, time: float):
    c = scene.to_c()
    out = (DpInjection * max(2 * len(scene.sources), 1))()
    _check(lib().or_scene_injections(byref(c), float(time), out, len(out)))
    return list(out)[: 2 * len(scene.sources)]


def injection_weight(scene: Scene, injection: DpInjection, coord) -> float:
    c = scene.to_c()
    cc = (c_uint32 * 3)(*[int(v) for v in coord])
    return float(lib().or_injection_weight(byref(c), byref(injection), cc))


def material_indices(scene: Scene) -> np.ndarray:
    c = scene.to_c()
    extent, _, _ = grid_resolve(scene.grid)
    out = np.zeros(extent[0] * extent[1] * extent[2], dtype=np.uint32)
    _check(lib().or_scene_material_indices(byref(c), out.ctypes.data_as(POINTER(c_uint32)), out.size))
    return out


class Energy:
    def __init__(self, electric: float, magnetic: float):
        self.electric = electric
        self.magnetic = magnetic

    def total(self) -> float:
        return self.electric + self.magnetic


class Simulation:
    """cpu::Simulation (src/cpu.rs:64-760), restated in C."""

    def __init__(self, scene: Scene, threads: int = 0):
        self._scene_c = scene.to_c()
        handle = lib().or_sim_create(byref(self._scene_c))
        if not handle:
            raise OracleError(_err())
        self._handle = handle
        lib().or_sim_set_threads(handle, threads)
        extent = (c_uint32 * 3)()
        lib().or_sim_extent(handle, extent)
        self.extent = tuple(int(v) for v in extent)
        self.cells = int(lib().or_sim_cells(handle))
        self.grid = scene.grid

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            lib().or_sim_destroy(handle)
            self._handle = None

    def set_threads(self, threads: int) -> None:
        lib().or_sim_set_threads(self._handle, threads)

    def step_count(self) -> int:
        return int(lib().or_sim_step_count(self._handle))

    def time(self) -> float:
        return float(lib().or_sim_time(self._handle))

    def advance(self) -> None:
        lib().or_sim_advance(self._handle, 1)

    def advance_by(self, steps: int) -> None:
        lib().or_sim_advance(self._handle, int(steps))

    def is_reversible(self) -> bool:
        return bool(lib().or_sim_is_reversible(self._handle))

    def reverse_by(self, steps: int) -> None:
        status = lib().or_sim_reverse(self._handle, int(steps))
        if status != 0:
            raise OracleError(_err())

    def reset(self) -> None:
        lib().or_sim_reset(self._handle)

    def set_accumulate_intensity(self, on: bool) -> None:
        lib().or_sim_set_accumulate_intensity(self._handle, 1 if on else 0)

    def accumulated_steps(self) -> int:
        return int(lib().or_sim_accumulated_steps(self._handle))

    def _view(self, ptr, count, dtype):
        return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)

    @property
    def electric_all(self) -> np.ndarray:
        """Writable view of E, component-major (electric_mut)."""
        return np.ctypeslib.as_array(lib().or_sim_electric(self._handle), shape=(3 * self.cells,))

    @property
    def magnetic_all(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_sim_magnetic(self._handle), shape=(3 * self.cells,))

    def electric(self, axis) -> np.ndarray:
        a = int(axis)
        return self.electric_all[a * self.cells:(a + 1) * self.cells]

    def magnetic(self, axis) -> np.ndarray:
        a = int(axis)
        return self.magnetic_all[a * self.cells:(a + 1) * self.cells]

    def intensity(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_sim_intensity(self._handle), shape=(self.cells,))

    # -- the f64 track (the same update in double, for error statements) --
    def f64_begin(self) -> None:
        _check(lib().or_sim_f64_begin(self._handle))

    def f64_advance_by(self, steps: int) -> None:
        _check(lib().or_sim_f64_advance(self._handle, steps))

    def electric_f64(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_sim_electric_f64(self._handle), shape=(3 * self.cells,))

    def magnetic_f64(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_sim_magnetic_f64(self._handle), shape=(3 * self.cells,))

    def intensity_f64(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().or_sim_intensity_f64(self._handle), shape=(self.cells,))

    def restore_fields(self, electric, magnetic, step: int) -> None:
        e = np.ascontiguousarray(electric, dtype=np.float32)
        h = np.ascontiguousarray(magnetic, dtype=np.float32)
        status = lib().or_sim_restore(self._handle, e.ctypes.data_as(POINTER(c_float)),
                                      h.ctypes.data_as(POINTER(c_float)), e.size, int(step))
        if status != 0:
            raise OracleError(_err())

    def snapshot(self):
        from diaphane_amd.timeline import Snapshot

        return Snapshot(self.step_count(), self.electric_all.copy(), self.magnetic_all.copy())

    def restore(self, snapshot) -> None:
        self.restore_fields(snapshot.electric, snapshot.magnetic, snapshot.step)

    def energy(self) -> Energy:
        ue, uh = c_double(), c_double()
        lib().or_sim_energy(self._handle, byref(ue), byref(uh))
        return Energy(ue.value, uh.value)

    def is_finite(self) -> bool:
        return bool(lib().or_sim_is_finite(self._handle))

    def sample_electric(self, axis, coord) -> float:
        nx, ny, _ = self.extent
        return float(self.electric(axis)[(coord[2] * ny + coord[1]) * nx + coord[0]])
