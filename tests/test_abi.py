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
    # The engine answers whether a device exists: dp_engine_create opens one
    # or fails with DP_ERR_DEVICE. A device index past the last one fails
    # the same way on any machine.
    from diaphane_amd.gpu import Context

    with pytest.raises(_lib.DiaphaneError) as caught:
        Context(4096)
    assert caught.value.code == _lib.DP_ERR_DEVICE
    try:
        context = Context(0)
    except _lib.DiaphaneError as error:
        assert error.code == _lib.DP_ERR_DEVICE, str(error)
        return
    assert context.arch().startswith("gfx")


def header_enums():
    values = {}
    for header in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for body in re.findall(r"enum\s*\{([^}]*)\}", text):
            for name, value in re.findall(r"\b(DP_[A-Z0-9_]+)\s*=\s*(-?\d+)", body):
                values[name] = int(value)
    return values


def test_python_constants_match_the_header():
    from diaphane_amd import batch, gpu

    enums = header_enums()
    pairs = {
        "DP_KERNEL_FUSED": gpu.KERNEL_FUSED, "DP_KERNEL_TWO_PASS": gpu.KERNEL_TWO_PASS,
        "DP_KERNEL_FUSED_OVERLAP": gpu.KERNEL_FUSED_OVERLAP,
        "DP_TILE_XCD_REMAP": gpu.TILE_XCD_REMAP, "DP_TILE_NONTEMPORAL": gpu.TILE_NONTEMPORAL,
        "DP_TILE_BUFFER": gpu.TILE_BUFFER, "DP_TILE_TWO_STEP": gpu.TILE_TWO_STEP, "DP_TILE_STACK": gpu.TILE_STACK,
        "DP_BUFFER_ELECTRIC": gpu.BUFFER_ELECTRIC, "DP_BUFFER_MAGNETIC": gpu.BUFFER_MAGNETIC,
        "DP_BUFFER_INTENSITY": gpu.BUFFER_INTENSITY, "DP_BUFFER_MATERIAL_INDEX": gpu.BUFFER_MATERIAL_INDEX,
        "DP_BUFFER_INTENSITY_F32": gpu.BUFFER_INTENSITY_F32, "DP_BUFFER_LOOKUP": gpu.BUFFER_LOOKUP,
        "DP_BATCH_WAVE_PER_TILE": batch.WAVE_PER_TILE,
    }
    for name, value in pairs.items():
        assert enums.get(name) == value, f"{name}: header {enums.get(name)}, binding {value}"
