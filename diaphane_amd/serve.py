"""Serving: a request loop with continuous batching over the device.

The reference has one ``gpu::Simulation`` per caller (src/gpu.rs:190). A
:class:`Server` accepts solves (a scene and a step count) from any thread,
admits them by a device-memory budget, builds their sims on host worker
threads while the device keeps stepping the others, and advances everything
admitted as one batch — a request that arrives while others run joins the
next launch. Finished sims stay resident until released.
"""
from __future__ import annotations

from ctypes import byref, c_uint64, c_void_p
from dataclasses import dataclass

from ._lib import DpRequestInfo, DpServerConfig, DpServerStats, check, lib
from .gpu import Context, Simulation
from .scene import Scene

QUEUED, PREPARING, READY, RUNNING, DONE = range(5)


@dataclass
class RequestInfo:
    state: int
    status: int
    steps: int
    step: int
    cells: int
    footprint_bytes: int
    submitted_s: float
    admitted_s: float
    started_s: float
    completed_s: float

    @property
    def latency_s(self) -> float:
        return self.completed_s - self.submitted_s


class Server:
    """A running request loop. ``one_at_a_time`` runs each request alone,
    in arrival order, through the plain single-sim path (the baseline).
    Otherwise resident requests share one batch; each chunk steps them in
    order of least remaining work until ``chunk_cells`` cells are stepping
    (0: 128M; ``ALL_MEMBERS``: every member every chunk)."""

    ALL_MEMBERS = (1 << 64) - 1

    def __init__(self, context: Context, hbm_budget_bytes: int = 200 << 30, max_resident: int = 1024,
                 chunk_steps: int = 32, host_threads: int = 4, one_at_a_time: bool = False, rows: int = 0,
                 depth: int = 0, chunk_cells: int = 0):
        config = DpServerConfig(hbm_budget_bytes, max_resident, chunk_steps, host_threads, int(one_at_a_time),
                                rows, depth, chunk_cells)
        handle = c_void_p()
        check(lib().dp_server_create(context.handle, byref(config), byref(handle)))
        self._handle = handle
        self._grids = {}

    def close(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle:
            lib().dp_server_destroy(handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def submit(self, scene: Scene, steps: int) -> int:
        scene_c = scene.to_c()
        ticket = c_uint64()
        check(lib().dp_server_submit(self._handle, byref(scene_c), steps, byref(ticket)))
        self._grids[ticket.value] = scene.grid
        return ticket.value

    def info(self, ticket: int) -> RequestInfo:
        raw = DpRequestInfo()
        check(lib().dp_server_query(self._handle, ticket, byref(raw)))
        return RequestInfo(*(getattr(raw, name) for name, _ in DpRequestInfo._fields_))

    def wait(self, ticket: int, timeout_ms: int = 0) -> None:
        check(lib().dp_server_wait(self._handle, ticket, timeout_ms))

    def result(self, ticket: int) -> Simulation:
        """The finished request's sim (borrowed until :meth:`release`)."""
        handle = c_void_p()
        check(lib().dp_server_result(self._handle, ticket, byref(handle)))
        return Simulation.borrowed(handle, self._grids.get(ticket))

    def release(self, ticket: int) -> None:
        check(lib().dp_server_release(self._handle, ticket))
        self._grids.pop(ticket, None)

    def stats(self) -> dict:
        raw = DpServerStats()
        check(lib().dp_server_statistics(self._handle, byref(raw)))
        return {name: getattr(raw, name) for name, _ in DpServerStats._fields_}
