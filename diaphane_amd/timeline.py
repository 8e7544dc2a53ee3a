"""Keyframes and seeking (src/timeline.rs).

The state is a pure function of (scene, step), so a seek is "restore the
newest keyframe at or before the target and replay", falling back to a reset
and replay from zero. Works over anything with the Steppable surface:
``step_count``, ``advance_by``, ``reset``, ``snapshot``, ``restore``.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass
from typing import Deque, Optional

import numpy as np


@dataclass
class Snapshot:
    """Both fields in full and the step they belong to (src/timeline.rs:43-55)."""

    step: int
    electric: np.ndarray
    magnetic: np.ndarray

    def bytes(self) -> int:
        return int(self.electric.nbytes + self.magnetic.nbytes)


@dataclass(frozen=True)
class Seek:
    """What a seek had to do (src/timeline.rs:85-110)."""

    kind: str  # "unchanged", "forward", "from_keyframe", "from_start"
    replayed: int = 0
    at: Optional[int] = None


class Timeline:
    """A ring of keyframes covering the recent past (src/timeline.rs:76-212)."""

    def __init__(self, interval: int, capacity: int):
        if interval <= 0:
            raise ValueError("keyframe interval must be positive")
        if capacity <= 0:
            raise ValueError("a timeline needs room for at least one keyframe")
        self.interval = interval
        self.capacity = capacity
        self.keyframes: Deque[Snapshot] = deque()

    def observe(self, simulation) -> None:
        step = simulation.step_count()
        if self.keyframes:
            last = self.keyframes[-1]
            if step <= last.step:
                return
            newest = last.step // self.interval
        else:
            newest = 0
        if step // self.interval <= newest:
            return
        if len(self.keyframes) == self.capacity:
            self.keyframes.popleft()
        self.keyframes.append(simulation.snapshot())

    def clear(self) -> None:
        self.keyframes.clear()

    def earliest(self) -> Optional[int]:
        return self.keyframes[0].step if self.keyframes else None

    def keyframe_count(self) -> int:
        return len(self.keyframes)

    def bytes(self) -> int:
        return sum(frame.bytes() for frame in self.keyframes)

    def seek(self, simulation, target: int) -> Seek:
        current = simulation.step_count()
        if target == current:
            return Seek("unchanged")
        if target > current:
            simulation.advance_by(target - current)
            return Seek("forward", target - current)
        for frame in reversed(self.keyframes):
            if frame.step <= target:
                simulation.restore(frame)
                simulation.advance_by(target - frame.step)
                return Seek("from_keyframe", target - frame.step, frame.step)
        simulation.reset()
        simulation.advance_by(target)
        return Seek("from_start", target)


class DeviceTimeline:
    """The same ring with its keyframes in device memory (dp_timeline): the
    snapshots never cross PCIe, and recording or restoring one is a
    device-to-device copy on the sim's stream. Works with a
    ``diaphane_amd.gpu.Simulation``."""

    _KINDS = {0: "unchanged", 1: "forward", 2: "from_keyframe", 3: "from_start"}

    def __init__(self, interval: int, capacity: int):
        import ctypes

        from ._lib import check, lib

        if interval <= 0:
            raise ValueError("keyframe interval must be positive")
        if capacity <= 0:
            raise ValueError("a timeline needs room for at least one keyframe")
        handle = ctypes.c_void_p()
        check(lib().dp_timeline_create(interval, capacity, ctypes.byref(handle)))
        self._handle = handle
        self.interval = interval
        self.capacity = capacity

    def close(self) -> None:
        from ._lib import lib

        handle = getattr(self, "_handle", None)
        if handle:
            lib().dp_timeline_destroy(handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def observe(self, simulation) -> None:
        from ._lib import check, lib

        check(lib().dp_timeline_observe(self._handle, simulation._handle))

    def seek(self, simulation, target: int) -> Seek:
        import ctypes

        from ._lib import check, lib

        kind, replayed, at = ctypes.c_int(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().dp_timeline_seek(self._handle, simulation._handle, target, ctypes.byref(kind),
                                     ctypes.byref(replayed), ctypes.byref(at)))
        name = self._KINDS[kind.value]
        return Seek(name, replayed.value, at.value if name == "from_keyframe" else None)

    def clear(self) -> None:
        from ._lib import check, lib

        check(lib().dp_timeline_clear(self._handle))

    def _info(self):
        import ctypes

        from ._lib import check, lib

        count, earliest, size = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().dp_timeline_info(self._handle, ctypes.byref(count), ctypes.byref(earliest), ctypes.byref(size)))
        return count.value, earliest.value, size.value

    def keyframe_count(self) -> int:
        return self._info()[0]

    def earliest(self) -> Optional[int]:
        count, earliest, _ = self._info()
        return earliest if count else None

    def bytes(self) -> int:
        return self._info()[2]
