"""diaphane_amd: an MI355X (gfx950) engine for diaphane's 3D Yee FDTD solver.

The scene layer mirrors kvark/diaphane's public types; ``gpu.Simulation``
is the device solver behind the C ABI in ``include/diaphane_amd.h``.
"""
from .scene import (
    AXES,
    COURANT_LIMIT,
    SPEED_OF_LIGHT,
    VACUUM_PERMEABILITY,
    VACUUM_PERMITTIVITY,
    Axis,
    Boundary,
    Extent,
    Grid,
    Material,
    MaterialTable,
    Refinement,
    Scene,
    Shape,
    Source,
    Waveform,
)
from .timeline import Seek, Snapshot, Timeline
from ._lib import DiaphaneError

__all__ = [
    "AXES", "COURANT_LIMIT", "SPEED_OF_LIGHT", "VACUUM_PERMEABILITY", "VACUUM_PERMITTIVITY",
    "Axis", "Boundary", "Extent", "Grid", "Material", "MaterialTable", "Refinement", "Scene",
    "Shape", "Source", "Waveform", "Seek", "Snapshot", "Timeline", "DiaphaneError",
]
