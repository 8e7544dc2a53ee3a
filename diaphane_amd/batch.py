"""Batched simulations: many independent grids advanced by one launch per step.

The reference runs one ``gpu::Simulation`` per scene (src/gpu.rs:159-484).
A :class:`Batch` advances several :class:`~diaphane_amd.gpu.Simulation`
objects together through ``dp_batch_*``; each member stays an ordinary
simulation for reading, writing, resetting and energy, and its fields are
bit-identical to the same simulation advanced alone.
"""
from __future__ import annotations

from ctypes import byref, c_float, c_int, c_void_p

from ._lib import check, lib
from .gpu import Context, Simulation

WAVE_PER_TILE = 16  # one wave per tile instead of the persistent grid


class Batch:
    """A set of simulations advanced together on one stream."""

    TIMEOUT_MS = 120_000

    def __init__(self, context: Context):
        self._context = context
        handle = c_void_p()
        check(lib().dp_batch_create(context.handle, byref(handle)))
        self._handle = handle
        self._members: dict[int, Simulation] = {}

    def close(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle:
            lib().dp_batch_destroy(handle)
            self._handle = None
            self._members.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, sim: Simulation, target: int | None = None) -> None:
        check(lib().dp_batch_add(self._handle, sim._handle))
        self._members[id(sim)] = sim
        if target is not None:
            self.set_target(sim, target)

    def remove(self, sim: Simulation) -> None:
        check(lib().dp_batch_remove(self._handle, sim._handle))
        self._members.pop(id(sim), None)

    def set_target(self, sim: Simulation, step: int) -> None:
        check(lib().dp_batch_set_target(self._handle, sim._handle, step))

    def advance_by(self, steps: int) -> None:
        check(lib().dp_batch_advance(self._handle, steps))

    def wait(self, timeout_ms: int = 0) -> None:
        check(lib().dp_batch_wait(self._handle, timeout_ms or self.TIMEOUT_MS))

    def set_graph_steps(self, steps: int) -> None:
        check(lib().dp_batch_set_graph_steps(self._handle, steps))

    def set_tile(self, rows: int = 0, depth: int = 0, flags: int = -1) -> None:
        check(lib().dp_batch_set_tile(self._handle, rows, depth, flags))

    def autotune(self, trials: int = 5) -> tuple[int, int, float]:
        """Times every rows × depth × form candidate (one or two steps per
        launch); keeps and returns the fastest as (rows, depth, ms per step).
        Member state is unchanged."""
        rows, depth, ms = c_int(), c_int(), c_float()
        check(lib().dp_batch_autotune(self._handle, trials, byref(rows), byref(depth), byref(ms)))
        return rows.value, depth.value, ms.value

    def tiling(self) -> dict:
        rows, depth, flags = c_int(), c_int(), c_int()
        check(lib().dp_batch_tiling(self._handle, byref(rows), byref(depth), byref(flags)))
        return {"rows": rows.value, "depth": depth.value, "wave_per_tile": bool(flags.value & WAVE_PER_TILE),
                "buffer": bool(flags.value & 4), "nontemporal": bool(flags.value & 2),
                "two_step": bool(flags.value & 8),
                "stack_waves": ((flags.value >> 16) & 0xF or 4) if flags.value & 32 else 0,
                "xcd_group": (((flags.value >> 8) & 0xFF) or 8) if flags.value & 1 else 0}

    def __len__(self) -> int:
        return int(lib().dp_batch_size(self._handle))

    def cells(self) -> int:
        return int(lib().dp_batch_cells(self._handle))

    def launch_count(self) -> int:
        return int(lib().dp_batch_launch_count(self._handle))

    def record_event(self, slot: int) -> None:
        check(lib().dp_batch_record_event(self._handle, slot))

    def elapsed_ms(self, start: int, end: int) -> float:
        ms = c_float()
        check(lib().dp_batch_elapsed_ms(self._handle, start, end, byref(ms)))
        return float(ms.value)
