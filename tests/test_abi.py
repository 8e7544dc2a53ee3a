"""The C ABI: the library loads and exports every entry point the public
header declares; host-only entry points behave without a device."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from diaphane_amd import _lib
from diaphane_amd import Axis, Extent, Grid, Scene

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(dp_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_engine_surface():
    names = declared_functions()
    for required in ("dp_sim_create", "dp_sim_advance", "dp_sim_reverse", "dp_sim_read_electric",
                     "dp_sim_write_fields", "dp_sim_energy", "dp_sim_set_accumulate_intensity",
                     "dp_sim_device_ptr", "dp_grid_resolve", "dp_scene_material_indices"):
        assert required in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports_every_declared_symbol(name):
    library = ctypes.CDLL(str(_lib.LIBRARY_PATH))
    assert hasattr(library, name), f"{name} is declared in include/ but not exported"


def test_errors_come_back_as_status_codes():
    grid = Grid.new(Extent.cube(16), 1e-3)
    grid.courant = 0.9
    with pytest.raises(_lib.DiaphaneError, match="Courant number"):
        grid.validate()
    assert _lib.lib().dp_version().decode().startswith("diaphane_amd")


def test_engine_creation_without_a_device_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from diaphane_amd.gpu import Context

    with pytest.raises(_lib.DiaphaneError):
        Context(0)
