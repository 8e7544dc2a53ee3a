"""Loads libdiaphane_amd.so and declares its C ABI (include/diaphane_amd.h).

The library is built in-tree by ``diaphane_amd.build``. There is no fallback:
if the shared object is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

from .scene import DpGridSpec, DpInjection, DpScene


class DpServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", c_uint64), ("max_resident", c_uint32), ("chunk_steps", c_uint32),
                ("host_threads", c_uint32), ("one_at_a_time", ctypes.c_int32), ("rows", ctypes.c_int32),
                ("depth", ctypes.c_int32), ("chunk_cells", c_uint64)]


class DpRequestInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("steps", c_uint64), ("step", c_uint64),
                ("cells", c_uint64), ("footprint_bytes", c_uint64), ("submitted_s", c_double),
                ("admitted_s", c_double), ("started_s", c_double), ("completed_s", c_double)]


class DpServerStats(ctypes.Structure):
    _fields_ = [("queued", c_uint32), ("preparing", c_uint32), ("ready", c_uint32), ("running", c_uint32),
                ("used_bytes", c_uint64), ("budget_bytes", c_uint64), ("chunks", c_uint64), ("busy_s", c_double),
                ("uptime_s", c_double), ("kernel_s", c_double), ("cell_updates", c_uint64)]

LIBRARY_PATH = Path(__file__).resolve().parent / "libdiaphane_amd.so"

DP_OK = 0
DP_ERR_INVALID = -1
DP_ERR_LOSSY = -2
DP_ERR_PAST_START = -3
DP_ERR_SIZE = -4
DP_ERR_TIMEOUT = -5
DP_ERR_DEVICE = -6
DP_ERR_UNSUPPORTED = -7
DP_ERR_BUDGET = -8


class DiaphaneError(RuntimeError):
    """A failed call into the engine; ``code`` is the DP_ERR_* status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


_lib = None


def _declare(lib: ctypes.CDLL) -> None:
    sim = c_void_p
    table = {
        "dp_last_error": (c_char_p, []),
        "dp_version": (c_char_p, []),
        "dp_host_alloc": (c_int, [c_size_t, POINTER(c_void_p)]),
        "dp_host_free": (None, [c_void_p]),
        "dp_grid_resolve": (c_int, [POINTER(DpGridSpec), POINTER(c_uint32), POINTER(c_float), c_size_t,
                                    POINTER(c_float)]),
        "dp_grid_cell_lookup": (c_int, [POINTER(DpGridSpec), c_uint32, POINTER(c_float), c_size_t]),
        "dp_grid_lookup_samples": (c_int, [POINTER(DpGridSpec), POINTER(c_uint32)]),
        "dp_scene_validate": (c_int, [POINTER(DpScene)]),
        "dp_scene_from_ron": (c_int, [c_char_p, POINTER(c_void_p)]),
        "dp_scene_file_scene": (POINTER(DpScene), [c_void_p]),
        "dp_scene_file_destroy": (None, [c_void_p]),
        "dp_scene_to_ron": (c_int, [POINTER(DpScene), c_char_p, c_size_t, POINTER(c_size_t)]),
        "dp_scene_material_indices": (c_int, [POINTER(DpScene), POINTER(c_uint32), c_size_t]),
        "dp_scene_injections": (c_int, [POINTER(DpScene), c_float, POINTER(DpInjection), c_size_t]),
        "dp_scene_tables": (c_int, [POINTER(DpScene), POINTER(c_float), c_size_t, POINTER(c_float), c_size_t]),
        "dp_engine_create": (c_int, [c_int, POINTER(c_void_p)]),
        "dp_engine_destroy": (None, [c_void_p]),
        "dp_engine_arch": (c_int, [c_void_p, ctypes.c_char_p, c_size_t]),
        "dp_sim_create": (c_int, [c_void_p, POINTER(DpScene), POINTER(c_void_p)]),
        "dp_sim_destroy": (None, [sim]),
        "dp_sim_extent": (c_int, [sim, POINTER(c_uint32)]),
        "dp_sim_step_count": (c_uint64, [sim]),
        "dp_sim_time": (c_float, [sim]),
        "dp_sim_time_step": (c_float, [sim]),
        "dp_sim_field_bytes": (c_uint64, [sim]),
        "dp_sim_bytes_per_step": (c_uint64, [sim]),
        "dp_sim_is_reversible": (c_int, [sim]),
        "dp_sim_advance": (c_int, [sim, c_uint64]),
        "dp_sim_reverse": (c_int, [sim, c_uint64]),
        "dp_sim_wait": (c_int, [sim, c_uint32]),
        "dp_sim_reset": (c_int, [sim]),
        "dp_sim_set_accumulate_intensity": (c_int, [sim, c_int]),
        "dp_sim_accumulated_steps": (c_uint64, [sim]),
        "dp_sim_read_electric": (c_int, [sim, POINTER(c_float), c_size_t]),
        "dp_sim_read_magnetic": (c_int, [sim, POINTER(c_float), c_size_t]),
        "dp_sim_begin_read": (c_int, [sim, c_int, POINTER(c_float), c_size_t]),
        "dp_sim_wait_reads": (c_int, [sim, c_uint32, c_uint32]),
        "dp_sim_finish_read": (c_int, [sim, c_uint32]),
        "dp_sim_read_intensity": (c_int, [sim, POINTER(c_float), c_size_t]),
        "dp_sim_read_intensity_f64": (c_int, [sim, POINTER(c_double), c_size_t]),
        "dp_sim_write_fields": (c_int, [sim, POINTER(c_float), POINTER(c_float), c_size_t, c_uint64]),
        "dp_sim_energy": (c_int, [sim, POINTER(c_double), POINTER(c_double)]),
        "dp_sim_device_ptr": (c_int, [sim, c_int, POINTER(c_void_p)]),
        "dp_sim_set_kernel": (c_int, [sim, c_int]),
        "dp_sim_set_tile": (c_int, [sim, c_int, c_int, c_int]),
        "dp_sim_autotune": (c_int, [sim, c_uint32, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(c_float)]),
        "dp_sim_record_event": (c_int, [sim, c_int]),
        "dp_sim_elapsed_ms": (c_int, [sim, c_int, c_int, POINTER(c_float)]),
        "dp_sim_launch_count": (c_uint64, [sim]),
        "dp_sim_lookup_samples": (c_uint32, [sim]),
        "dp_sim_stream_probe": (c_int, [sim, c_int, POINTER(c_double)]),
        "dp_sim_tile_probe": (c_int, [sim, c_int, c_int, c_int, POINTER(c_double)]),
        "dp_sim_measure_peak": (c_int, [sim, POINTER(c_float)]),
        "dp_sim_measure_peak_noalloc": (c_int, [sim, POINTER(c_float)]),
        "dp_timeline_create": (c_int, [c_uint64, c_uint32, POINTER(c_void_p)]),
        "dp_timeline_destroy": (None, [c_void_p]),
        "dp_timeline_observe": (c_int, [c_void_p, sim]),
        "dp_timeline_seek": (c_int, [c_void_p, sim, c_uint64, POINTER(c_int), POINTER(c_uint64), POINTER(c_uint64)]),
        "dp_timeline_clear": (c_int, [c_void_p]),
        "dp_timeline_info": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint64), POINTER(c_uint64)]),
        "dp_batch_create": (c_int, [c_void_p, POINTER(c_void_p)]),
        "dp_batch_destroy": (None, [c_void_p]),
        "dp_batch_add": (c_int, [c_void_p, sim]),
        "dp_batch_remove": (c_int, [c_void_p, sim]),
        "dp_batch_set_target": (c_int, [c_void_p, sim, c_uint64]),
        "dp_batch_advance": (c_int, [c_void_p, c_uint64]),
        "dp_batch_wait": (c_int, [c_void_p, c_uint32]),
        "dp_batch_set_graph_steps": (c_int, [c_void_p, c_uint32]),
        "dp_batch_set_tile": (c_int, [c_void_p, c_int, c_int, c_int]),
        "dp_batch_autotune": (c_int, [c_void_p, c_uint32, POINTER(c_int), POINTER(c_int), POINTER(c_float)]),
        "dp_batch_tiling": (c_int, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "dp_batch_size": (c_uint32, [c_void_p]),
        "dp_batch_cells": (c_uint64, [c_void_p]),
        "dp_batch_launch_count": (c_uint64, [c_void_p]),
        "dp_batch_record_event": (c_int, [c_void_p, c_int]),
        "dp_batch_elapsed_ms": (c_int, [c_void_p, c_int, c_int, POINTER(c_float)]),
        "dp_server_create": (c_int, [c_void_p, POINTER(DpServerConfig), POINTER(c_void_p)]),
        "dp_server_destroy": (None, [c_void_p]),
        "dp_server_submit": (c_int, [c_void_p, POINTER(DpScene), c_uint64, POINTER(c_uint64)]),
        "dp_server_query": (c_int, [c_void_p, c_uint64, POINTER(DpRequestInfo)]),
        "dp_server_wait": (c_int, [c_void_p, c_uint64, c_uint32]),
        "dp_server_result": (c_int, [c_void_p, c_uint64, POINTER(c_void_p)]),
        "dp_server_release": (c_int, [c_void_p, c_uint64]),
        "dp_server_statistics": (c_int, [c_void_p, POINTER(DpServerStats)]),
    }
    for name, (restype, argtypes) in table.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes


def lib() -> ctypes.CDLL:
    """The loaded engine library (loaded on first use)."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise DiaphaneError(
                DP_ERR_UNSUPPORTED,
                f"{LIBRARY_PATH} is not built; run `python -m diaphane_amd.build` (hipcc, gfx950)",
            )
        loaded = ctypes.CDLL(str(LIBRARY_PATH))
        _declare(loaded)
        _lib = loaded
    return _lib


def last_error() -> str:
    message = lib().dp_last_error()
    return message.decode() if message else ""


def check(status: int) -> None:
    if status != DP_OK:
        raise DiaphaneError(status, last_error())
