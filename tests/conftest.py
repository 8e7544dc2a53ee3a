"""Shared fixtures. The engine library and the oracle are built in-tree if
they are missing (a build is a no-op when they already exist)."""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built engine")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def _ensure_built():
    from diaphane_amd import _lib

    if not _lib.LIBRARY_PATH.exists():
        from diaphane_amd import build

        build.build()
    import oracle

    if not oracle.LIBRARY_PATH.exists():
        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def context():
    """The HIP device; a GPU test fails (not skips) when it cannot open one."""
    from diaphane_amd.gpu import headless_context

    return headless_context()


def peak_of(values) -> float:
    return float(np.max(np.abs(values))) if np.size(values) else 0.0


def assert_field_parity(name, reference, candidate, steps, tolerance, guard=2e-3):
    """tests/parity.rs:51-86: per component, max|Δ| over the peak across all
    three components, with the absorbed-remnant guard."""
    peak = max(peak_of(r) for r in reference)
    assert peak > 0.0, f"{name} is identically zero after {steps} steps"
    assert peak > guard, f"{name} peak {peak:e} after {steps} steps is an absorbed remnant"
    worst = 0.0
    for axis, (host, device) in enumerate(zip(reference, candidate)):
        relative = float(np.max(np.abs(host.astype(np.float64) - device.astype(np.float64)))) / peak
        worst = max(worst, relative)
        assert relative < tolerance, (
            f"{name}[{axis}] diverged by {relative:e} of the {peak:e} field peak after {steps} steps")
    return worst
