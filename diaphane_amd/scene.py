"""The scene layer, mirroring the reference's public types.

``Axis``, ``Extent``, ``Refinement``, ``Grid`` (src/grid.rs), ``Material``,
``MaterialTable`` (src/material.rs), ``Boundary`` (src/boundary.rs),
``Waveform``, ``Source`` (src/source.rs), ``Shape`` and ``Scene``
(src/scene.rs). Values that the reference computes in f32 are computed in
numpy float32 here, in the same order, so a scene built in Python describes
bit-for-bit the same problem as one built in Rust.

These classes only describe a problem. Grid resolution, validation and
rasterization run in the C++ host library (``dp_grid_resolve``,
``dp_scene_validate``, ``dp_scene_material_indices``).
"""
from __future__ import annotations

import ctypes
import enum
import math
from ctypes import Structure, c_float, c_uint32
from dataclasses import dataclass, field, replace
from typing import List, Optional, Sequence, Tuple

import numpy as np

f32 = np.float32

SPEED_OF_LIGHT = f32(299792458.0)
VACUUM_PERMEABILITY = f32(1.2566371e-6)
VACUUM_PERMITTIVITY = f32(8.854188e-12)
VACUUM_IMPEDANCE = f32(376.73032)
COURANT_LIMIT = f32(0.57735026)
PI = f32(math.pi)


# ---------------------------------------------------------------------------
# C structures (include/diaphane_amd.h)
# ---------------------------------------------------------------------------
class DpRefinement(Structure):
    _fields_ = [("center", c_float * 3), ("size", c_float * 3), ("cell_size", c_float)]


class DpGridSpec(Structure):
    _fields_ = [
        ("extent", c_uint32 * 3),
        ("cell_size", c_float),
        ("courant", c_float),
        ("max_ratio", c_float),
        ("refinements", ctypes.POINTER(DpRefinement)),
        ("refinement_count", c_uint32),
    ]


class DpMaterial(Structure):
    _fields_ = [
        ("relative_permittivity", c_float),
        ("relative_permeability", c_float),
        ("conductivity", c_float),
        ("magnetic_loss", c_float),
    ]


class DpShape(Structure):
    _fields_ = [
        ("kind", c_uint32),
        ("material", c_uint32),
        ("center", c_float * 3),
        ("size", c_float * 3),
        ("radius", c_float),
        ("axis", c_uint32),
        ("offset", c_float),
        ("thickness", c_float),
    ]


class DpWaveform(Structure):
    _fields_ = [
        ("kind", c_uint32),
        ("frequency", c_float),
        ("cycles", c_float),
        ("delay", c_float),
        ("width", c_float),
    ]


class DpSource(Structure):
    _fields_ = [
        ("shape", c_uint32),
        ("polarization", c_uint32),
        ("at", c_float * 3),
        ("velocity", c_float * 3),
        ("axis", c_uint32),
        ("offset", c_float),
        ("waist", c_float),
        ("waveform", DpWaveform),
        ("amplitude", c_float),
    ]


class DpScene(Structure):
    _fields_ = [
        ("grid", DpGridSpec),
        ("boundary", c_uint32),
        ("absorber_thickness", c_uint32),
        ("target_reflection", c_float),
        ("materials", ctypes.POINTER(DpMaterial)),
        ("material_count", c_uint32),
        ("shapes", ctypes.POINTER(DpShape)),
        ("shape_count", c_uint32),
        ("sources", ctypes.POINTER(DpSource)),
        ("source_count", c_uint32),
    ]


class DpInjection(Structure):
    _fields_ = [
        ("origin", c_uint32 * 3),
        ("extent", c_uint32 * 3),
        ("component", c_uint32),
        ("magnetic", c_uint32),
        ("value", c_float),
        ("center", c_float * 3),
        ("inverse_waist_squared", c_float),
    ]


# ---------------------------------------------------------------------------
# Grid (src/grid.rs)
# ---------------------------------------------------------------------------
class Axis(enum.IntEnum):
    X = 0
    Y = 1
    Z = 2

    def index(self) -> int:
        return int(self)

    def next(self) -> "Axis":
        return Axis((int(self) + 1) % 3)

    def prev(self) -> "Axis":
        return Axis((int(self) + 2) % 3)


AXES = (Axis.X, Axis.Y, Axis.Z)


@dataclass(frozen=True)
class Extent:
    x: int
    y: int
    z: int

    @staticmethod
    def cube(side: int) -> "Extent":
        return Extent(side, side, side)

    def as_array(self) -> Tuple[int, int, int]:
        return (self.x, self.y, self.z)

    def total(self) -> int:
        return self.x * self.y * self.z

    def strides(self) -> Tuple[int, int, int]:
        return (1, self.x, self.x * self.y)

    def index(self, coord: Sequence[int]) -> int:
        return (coord[2] * self.y + coord[1]) * self.x + coord[0]


@dataclass(frozen=True)
class Refinement:
    center: Tuple[float, float, float]
    size: Tuple[float, float, float]
    cell_size: float

    def __post_init__(self):
        object.__setattr__(self, "center", tuple(float(f32(v)) for v in self.center))
        object.__setattr__(self, "size", tuple(float(f32(v)) for v in self.size))
        object.__setattr__(self, "cell_size", float(f32(self.cell_size)))

    @staticmethod
    def across(axis: Axis, offset: float, size: float, cell_size: float) -> "Refinement":
        center = [0.0, 0.0, 0.0]
        extent = [0.0, 0.0, 0.0]
        center[int(axis)] = offset
        extent[int(axis)] = size
        return Refinement(tuple(center), tuple(extent), cell_size)


def _round_half_away(value: float) -> int:
    return int(math.floor(value + 0.5)) if value >= 0 else -int(math.floor(-value + 0.5))


def extent_for(size: Sequence[float], resolution: float) -> Extent:
    """Grid::extent_for (src/grid.rs:458-464)."""
    if not (resolution > 0 and math.isfinite(resolution)):
        raise ValueError(f"resolution must be positive and finite, got {resolution}")
    cells = [max(_round_half_away(float(f32(m) * f32(resolution))), 2) for m in size]
    return Extent(*cells)


class Grid:
    """The discretization: a GridSpec recipe plus its resolved spacing."""

    COURANT_LIMIT = COURANT_LIMIT

    def __init__(
        self,
        base_extent: Extent,
        cell_size: float,
        courant: float = 0.5,
        refinements: Sequence[Refinement] = (),
        max_ratio: float = 1.15,
    ):
        self.base_extent = base_extent
        self.base_cell_size = float(f32(cell_size))
        self.refinements = list(refinements)
        self.max_ratio = float(f32(max_ratio))
        self._courant = float(f32(courant))
        self._resolved = None

    def _recipe(self):
        return (self.base_extent, self.base_cell_size, self._courant, tuple(self.refinements), self.max_ratio)

    def __eq__(self, other) -> bool:
        """Grids are equal when their recipes are (the resolved spacing follows)."""
        return isinstance(other, Grid) and self._recipe() == other._recipe()

    __hash__ = None

    # constructors (src/grid.rs:415-456)
    @staticmethod
    def new(extent: Extent, cell_size: float) -> "Grid":
        return Grid(extent, cell_size)

    @staticmethod
    def for_size(size: Sequence[float], resolution: float) -> "Grid":
        return Grid(extent_for(size, resolution), float(f32(1.0) / f32(resolution)))

    @staticmethod
    def graded(size: Sequence[float], resolution: float, refinements: Sequence[Refinement]) -> "Grid":
        return Grid.graded_at(size, resolution, refinements, 1.15)

    @staticmethod
    def graded_at(size, resolution, refinements, max_ratio) -> "Grid":
        return Grid(extent_for(size, resolution), float(f32(1.0) / f32(resolution)), 0.5,
                    refinements, max_ratio)

    def with_resolution(self, resolution: float) -> "Grid":
        return Grid(extent_for(self.size(), resolution), float(f32(1.0) / f32(resolution)), self.courant,
                    self.refinements, self.max_ratio)

    @property
    def courant(self) -> float:
        return self._courant

    @courant.setter
    def courant(self, value: float) -> None:
        self._courant = float(f32(value))
        self._resolved = None

    def spec(self) -> DpGridSpec:
        spec = DpGridSpec()
        spec.extent[:] = list(self.base_extent.as_array())
        spec.cell_size = self.base_cell_size
        spec.courant = self._courant
        spec.max_ratio = self.max_ratio
        refinements = (DpRefinement * max(len(self.refinements), 1))()
        for i, r in enumerate(self.refinements):
            refinements[i].center[:] = list(r.center)
            refinements[i].size[:] = list(r.size)
            refinements[i].cell_size = r.cell_size
        spec.refinements = refinements
        spec.refinement_count = len(self.refinements)
        spec._keep = refinements  # keep the array alive with the struct
        return spec

    def _resolve(self):
        if self._resolved is None:
            from ._lib import check, lib

            spec = self.spec()
            extent = (c_uint32 * 3)()
            check(lib().dp_grid_resolve(ctypes.byref(spec), extent, None, 0, None))
            n = int(extent[0]) + int(extent[1]) + int(extent[2])
            widths = np.zeros(n, dtype=np.float32)
            dt = c_float()
            check(lib().dp_grid_resolve(ctypes.byref(spec), extent,
                                        widths.ctypes.data_as(ctypes.POINTER(c_float)), n,
                                        ctypes.byref(dt)))
            ext = Extent(int(extent[0]), int(extent[1]), int(extent[2]))
            split = np.cumsum([0, ext.x, ext.y, ext.z])
            primary = [widths[split[a]:split[a + 1]].copy() for a in range(3)]
            self._resolved = (ext, primary, float(dt.value))
        return self._resolved

    @property
    def extent(self) -> Extent:
        return self._resolve()[0]

    def primary(self, axis: Axis) -> np.ndarray:
        return self._resolve()[1][int(axis)]

    def cell_lookup(self, samples: int) -> np.ndarray:
        """Grid::cell_lookup (src/grid.rs:635-646): 3 × samples fractional
        cell coordinates, x then y then z, as the engine uploads them."""
        from ._lib import check, lib

        out = np.zeros(3 * int(samples), dtype=np.float32)
        check(lib().dp_grid_cell_lookup(ctypes.byref(self.spec()), int(samples),
                                        out.ctypes.data_as(ctypes.POINTER(c_float)), out.size))
        return out

    def lookup_samples(self) -> int:
        """The sample count gpu::Simulation::new picks (src/gpu.rs:225-227)."""
        from ._lib import check, lib

        out = c_uint32()
        check(lib().dp_grid_lookup_samples(ctypes.byref(self.spec()), ctypes.byref(out)))
        return int(out.value)

    def dual(self, axis: Axis) -> np.ndarray:
        w = self.primary(axis)
        prev = np.concatenate([w[:1], w[:-1]])
        return (f32(0.5) * (prev + w)).astype(np.float32)

    def is_uniform(self) -> bool:
        return not self.refinements

    def time_step(self) -> float:
        return self._resolve()[2]

    def length(self, axis: Axis) -> float:
        """Spacing::length: the last corner, summed in f64 (src/grid.rs:172-189)."""
        w = self.primary(axis)
        if self.is_uniform():
            return float(f32(len(w) * float(w[0])))
        total = 0.0
        for value in w:
            total += float(value)
        return float(f32(total))

    def size(self) -> Tuple[float, float, float]:
        return tuple(self.length(a) for a in AXES)

    def finest(self) -> float:
        return float(min(self.primary(a).min() for a in AXES))

    def coarsest(self) -> float:
        return float(max(self.primary(a).max() for a in AXES))

    def worst_ratio(self) -> float:
        worst = f32(1.0)
        for a in AXES:
            w = self.primary(a)
            if len(w) > 1:
                r = np.maximum(w[1:] / w[:-1], w[:-1] / w[1:]).max()
                worst = max(worst, r)
        return float(worst)

    def frequency_for_resolution(self, cells: float) -> float:
        """src/grid.rs:742-744."""
        return float(SPEED_OF_LIGHT / (f32(cells) * f32(self.coarsest())))

    def cells_per_wavelength(self, frequency: float, refractive_index: float) -> float:
        return float(SPEED_OF_LIGHT / (f32(refractive_index) * f32(frequency) * f32(self.coarsest())))

    def validate(self) -> None:
        """Grid::validate (src/grid.rs:747-773): raises on an unstable grid."""
        self._resolved = None
        self._resolve()


# ---------------------------------------------------------------------------
# Materials (src/material.rs)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Material:
    relative_permittivity: float = 1.0
    relative_permeability: float = 1.0
    conductivity: float = 0.0
    magnetic_loss: float = 0.0

    def __post_init__(self):
        for name in ("relative_permittivity", "relative_permeability", "conductivity", "magnetic_loss"):
            object.__setattr__(self, name, float(f32(getattr(self, name))))

    @staticmethod
    def dielectric(relative_permittivity: float) -> "Material":
        return Material(relative_permittivity)

    @staticmethod
    def refractive(index: float) -> "Material":
        return Material(float(f32(index) * f32(index)))

    @staticmethod
    def matched_lossy(relative_permittivity: float, conductivity: float) -> "Material":
        eps = f32(relative_permittivity)
        sigma = f32(conductivity)
        return Material(float(eps), 1.0, float(sigma),
                        float(sigma * VACUUM_PERMEABILITY / (VACUUM_PERMITTIVITY * eps)))

    def refractive_index(self) -> float:
        return float(np.sqrt(f32(self.relative_permittivity) * f32(self.relative_permeability)))


Material.VACUUM = Material()
Material.PERFECT_CONDUCTOR = Material(1.0, 1.0, float("inf"), 0.0)


class MaterialTable:
    VACUUM = 0

    def __init__(self, materials: Optional[Sequence[Material]] = None):
        self.materials: List[Material] = list(materials) if materials is not None else [Material.VACUUM]

    def push(self, material: Material) -> int:
        for i, m in enumerate(self.materials):
            if m == material:
                return i
        self.materials.append(material)
        return len(self.materials) - 1

    def get(self, index: int) -> Material:
        return self.materials[index]

    def __len__(self) -> int:
        return len(self.materials)

    def copy(self) -> "MaterialTable":
        return MaterialTable(self.materials)

    def __eq__(self, other) -> bool:
        return isinstance(other, MaterialTable) and self.materials == other.materials

    __hash__ = None


# ---------------------------------------------------------------------------
# Boundary (src/boundary.rs)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Boundary:
    kind: str  # "pec" or "absorbing"
    thickness: int = 0
    target_reflection: float = 1.0

    @staticmethod
    def absorbing(thickness: int, target_reflection: float) -> "Boundary":
        return Boundary("absorbing", thickness, float(f32(target_reflection)))


Boundary.PEC = Boundary("pec")
Boundary.DEFAULT = Boundary.absorbing(10, 1e-6)


# ---------------------------------------------------------------------------
# Sources (src/source.rs)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Waveform:
    kind: int
    frequency: float = 0.0
    cycles: float = 0.0
    delay: float = 0.0
    width: float = 0.0

    RICKER = 0
    GAUSSIAN_PULSE = 1
    CONTINUOUS_WAVE = 2
    DIFF_GAUSSIAN = 3

    @staticmethod
    def Ricker(peak_frequency: float, delay: float) -> "Waveform":
        return Waveform(0, float(f32(peak_frequency)), 0.0, float(f32(delay)))

    @staticmethod
    def ricker(peak_frequency: float) -> "Waveform":
        f = f32(peak_frequency)
        return Waveform(0, float(f), 0.0, float(f32(1.4) / f))

    @staticmethod
    def GaussianPulse(center_frequency: float, duration_cycles: float, delay: float) -> "Waveform":
        return Waveform(1, float(f32(center_frequency)), float(f32(duration_cycles)), float(f32(delay)))

    @staticmethod
    def gaussian_pulse(center_frequency: float, duration_cycles: float) -> "Waveform":
        f, c = f32(center_frequency), f32(duration_cycles)
        return Waveform(1, float(f), float(c), float(f32(2.0) * c / f))

    @staticmethod
    def ContinuousWave(frequency: float, ramp_cycles: float) -> "Waveform":
        return Waveform(2, float(f32(frequency)), float(f32(ramp_cycles)))

    @staticmethod
    def DifferentiatedGaussian(width: float, delay: float) -> "Waveform":
        """−u·exp(½(1−u²)), u = (t − delay)/width; width is σ in seconds."""
        return Waveform(3, 0.0, 0.0, float(f32(delay)), float(f32(width)))

    def dominant_frequency(self) -> float:
        if self.kind == 3:
            return float(f32(1.0) / (f32(2.0) * PI * f32(self.width)))
        return self.frequency

    def to_c(self) -> DpWaveform:
        return DpWaveform(self.kind, self.frequency, self.cycles, self.delay, self.width)


@dataclass(frozen=True)
class Source:
    shape: int  # 0 point, 1 sheet, 2 plane wave
    polarization: Axis
    waveform: Waveform
    amplitude: float = 1.0
    at: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    axis: Axis = Axis.X
    offset: float = 0.0
    waist: float = 0.0

    POINT = 0
    SHEET = 1
    PLANE_WAVE = 2

    @staticmethod
    def point(at, polarization: Axis, waveform: Waveform) -> "Source":
        return Source(0, Axis(polarization), waveform, 1.0, tuple(float(f32(v)) for v in at))

    @staticmethod
    def point_moving(at, velocity, polarization: Axis, waveform: Waveform) -> "Source":
        return Source(0, Axis(polarization), waveform, 1.0, tuple(float(f32(v)) for v in at),
                      tuple(float(f32(v)) for v in velocity))

    @staticmethod
    def sheet(axis: Axis, offset: float, waist: float, polarization: Axis, waveform: Waveform) -> "Source":
        return Source(1, Axis(polarization), waveform, 1.0, axis=Axis(axis), offset=float(f32(offset)),
                      waist=float(f32(waist)))

    @staticmethod
    def plane_wave(axis: Axis, offset: float, polarization: Axis, waveform: Waveform) -> "Source":
        return Source(2, Axis(polarization), waveform, 1.0, axis=Axis(axis), offset=float(f32(offset)))

    def with_amplitude(self, amplitude: float) -> "Source":
        return replace(self, amplitude=float(f32(amplitude)))

    def position(self) -> Tuple[float, float, float]:
        if self.shape == 0:
            return self.at
        p = [0.0, 0.0, 0.0]
        p[int(self.axis)] = self.offset
        return tuple(p)

    def to_c(self) -> DpSource:
        s = DpSource()
        s.shape = self.shape
        s.polarization = int(self.polarization)
        s.at[:] = list(self.at)
        s.velocity[:] = list(self.velocity)
        s.axis = int(self.axis)
        s.offset = self.offset
        s.waist = self.waist
        s.waveform = self.waveform.to_c()
        s.amplitude = self.amplitude
        return s


# ---------------------------------------------------------------------------
# Shapes and scenes (src/scene.rs)
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Shape:
    kind: int  # 0 block, 1 sphere, 2 slab
    material: int
    center: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    size: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    radius: float = 0.0
    axis: Axis = Axis.X
    offset: float = 0.0
    thickness: float = 0.0

    @staticmethod
    def Block(center, size, material: int) -> "Shape":
        return Shape(0, material, tuple(float(f32(v)) for v in center), tuple(float(f32(v)) for v in size))

    @staticmethod
    def Sphere(center, radius: float, material: int) -> "Shape":
        return Shape(1, material, tuple(float(f32(v)) for v in center), radius=float(f32(radius)))

    @staticmethod
    def Slab(axis: Axis, offset: float, thickness: float, material: int) -> "Shape":
        return Shape(2, material, axis=Axis(axis), offset=float(f32(offset)), thickness=float(f32(thickness)))

    def to_c(self) -> DpShape:
        s = DpShape()
        s.kind = self.kind
        s.material = self.material
        s.center[:] = list(self.center)
        s.size[:] = list(self.size)
        s.radius = self.radius
        s.axis = int(self.axis)
        s.offset = self.offset
        s.thickness = self.thickness
        return s


PRESET_CELL_SIZE = 1e-3


@dataclass
class Scene:
    grid: Grid
    boundary: Boundary = Boundary.DEFAULT
    materials: MaterialTable = field(default_factory=MaterialTable)
    shapes: List[Shape] = field(default_factory=list)
    sources: List[Source] = field(default_factory=list)

    @staticmethod
    def sized(size, resolution: float) -> "Scene":
        return Scene(Grid.for_size(size, resolution))

    @staticmethod
    def empty(extent: Extent, cell_size: float) -> "Scene":
        return Scene(Grid.new(extent, cell_size))

    @staticmethod
    def on_grid(grid: Grid) -> "Scene":
        return Scene(grid)

    def with_resolution(self, resolution: float) -> "Scene":
        """Scene::with_resolution (src/scene.rs:158-163): the same scene on a
        grid of `resolution` cells per metre."""
        return Scene(self.grid.with_resolution(resolution), self.boundary, self.materials.copy(),
                     list(self.shapes), list(self.sources))

    # scene files (src/scene.rs:467-498)
    @staticmethod
    def from_ron(text: str) -> "Scene":
        """Scene::from_ron: parses the RON text format (not validated).
        Raises ValueError with the line and column of a syntax error."""
        from ._lib import DiaphaneError, lib

        handle = ctypes.c_void_p()
        status = lib().dp_scene_from_ron(text.encode("utf-8"), ctypes.byref(handle))
        if status != 0:
            from ._lib import last_error
            raise ValueError(last_error())
        try:
            return Scene.from_c(lib().dp_scene_file_scene(handle).contents)
        finally:
            lib().dp_scene_file_destroy(handle)

    def to_ron(self) -> str:
        """Scene::to_ron: text that from_ron reads back to an equal scene."""
        from ._lib import check, lib

        c = self.to_c()
        length = ctypes.c_size_t()
        check(lib().dp_scene_to_ron(ctypes.byref(c), None, 0, ctypes.byref(length)))
        buffer = ctypes.create_string_buffer(length.value + 1)
        check(lib().dp_scene_to_ron(ctypes.byref(c), buffer, len(buffer), ctypes.byref(length)))
        return buffer.value.decode("utf-8")

    @staticmethod
    def load(path) -> "Scene":
        """Scene::load: reads a scene file; errors name the file."""
        with open(path, encoding="utf-8") as handle:
            text = handle.read()
        try:
            return Scene.from_ron(text)
        except ValueError as error:
            raise ValueError(f"{path}: {error}") from None

    def save(self, path) -> None:
        with open(path, "w", encoding="utf-8") as handle:
            handle.write(self.to_ron())

    @staticmethod
    def from_c(c: "DpScene") -> "Scene":
        """A Python scene holding the values of a C dp_scene."""
        g = c.grid
        refinements = [Refinement(tuple(g.refinements[k].center), tuple(g.refinements[k].size),
                                  g.refinements[k].cell_size) for k in range(g.refinement_count)]
        grid = Grid(Extent(*g.extent), g.cell_size, g.courant, refinements, g.max_ratio)
        boundary = Boundary.PEC if c.boundary == 0 else Boundary.absorbing(c.absorber_thickness,
                                                                           c.target_reflection)
        materials = MaterialTable([Material(m.relative_permittivity, m.relative_permeability, m.conductivity,
                                            m.magnetic_loss) for m in c.materials[:c.material_count]])
        shapes = [Shape(s.kind, s.material, tuple(s.center), tuple(s.size), s.radius, Axis(s.axis), s.offset,
                        s.thickness) for s in c.shapes[:c.shape_count]]
        sources = []
        for s in c.sources[:c.source_count]:
            w = s.waveform
            sources.append(Source(s.shape, Axis(s.polarization), Waveform(w.kind, w.frequency, w.cycles, w.delay,
                                                                          w.width),
                                  s.amplitude, tuple(s.at), tuple(s.velocity), Axis(s.axis), s.offset, s.waist))
        return Scene(grid, boundary, materials, shapes, sources)

    def with_boundary(self, boundary: Boundary) -> "Scene":
        self.boundary = boundary
        return self

    def with_source(self, source: Source) -> "Scene":
        self.sources.append(source)
        return self

    def with_shape(self, shape: Shape) -> "Scene":
        self.shapes.append(shape)
        return self

    # presets (src/scene.rs:376-456)
    @staticmethod
    def photon(extent: Extent) -> "Scene":
        grid = Grid.new(extent, PRESET_CELL_SIZE)
        size = grid.size()
        frequency = grid.frequency_for_resolution(16.0)
        source = Source.sheet(Axis.X, float(f32(-0.3) * f32(size[0])),
                              float(f32(0.25) * min(f32(size[1]), f32(size[2]))), Axis.Z,
                              Waveform.gaussian_pulse(frequency, 2.0))
        return Scene(grid, Boundary.DEFAULT, MaterialTable(), [], [source])

    @staticmethod
    def cavity(extent: Extent) -> "Scene":
        return Scene.cavity_on(Grid.new(extent, PRESET_CELL_SIZE))

    @staticmethod
    def cavity_on(grid: Grid) -> "Scene":
        size = grid.size()
        frequency = grid.frequency_for_resolution(16.0)
        source = Source.point([float(-f32(size[0]) / f32(6.0)), 0.0, 0.0], Axis.Z, Waveform.ricker(frequency))
        return Scene(grid, Boundary.PEC, MaterialTable(), [], [source])

    @staticmethod
    def slab(extent: Extent, refractive_index: float) -> "Scene":
        scene = Scene.photon(extent)
        size = scene.grid.size()
        material = scene.materials.push(Material.refractive(refractive_index))
        scene.shapes.append(Shape.Slab(Axis.X, float(f32(size[0]) / f32(8.0)), float(f32(size[0]) / f32(4.0)),
                                       material))
        frequency = scene.grid.frequency_for_resolution(float(f32(16.0) * f32(refractive_index)))
        scene.sources = [replace(s, waveform=Waveform.gaussian_pulse(frequency, 2.0)) for s in scene.sources]
        return scene

    def is_reversible(self) -> bool:
        return self.boundary == Boundary.PEC and all(
            m.conductivity == 0.0 and m.magnetic_loss == 0.0 for m in self.materials.materials)

    def to_c(self) -> DpScene:
        scene = DpScene()
        spec = self.grid.spec()
        scene.grid = spec
        keep = [spec._keep]
        if self.boundary.kind == "pec":
            scene.boundary = 0
        else:
            scene.boundary = 1
            scene.absorber_thickness = self.boundary.thickness
            scene.target_reflection = self.boundary.target_reflection
        materials = (DpMaterial * max(len(self.materials), 1))()
        for i, m in enumerate(self.materials.materials):
            materials[i] = DpMaterial(m.relative_permittivity, m.relative_permeability, m.conductivity,
                                      m.magnetic_loss)
        shapes = (DpShape * max(len(self.shapes), 1))()
        for i, s in enumerate(self.shapes):
            shapes[i] = s.to_c()
        sources = (DpSource * max(len(self.sources), 1))()
        for i, s in enumerate(self.sources):
            sources[i] = s.to_c()
        scene.materials = materials
        scene.material_count = len(self.materials)
        scene.shapes = shapes
        scene.shape_count = len(self.shapes)
        scene.sources = sources
        scene.source_count = len(self.sources)
        keep += [materials, shapes, sources]
        scene._keep = keep
        return scene

    def validate(self) -> None:
        """Scene::validate (src/scene.rs:254-369); raises ValueError."""
        from ._lib import DiaphaneError, lib, last_error

        c = self.to_c()
        status = lib().dp_scene_validate(ctypes.byref(c))
        if status != 0:
            raise ValueError(last_error())

    def material_indices(self) -> np.ndarray:
        """Scene::material_indices (src/scene.rs:181-206), as u32."""
        from ._lib import check, lib

        c = self.to_c()
        out = np.zeros(self.grid.extent.total(), dtype=np.uint32)
        check(lib().dp_scene_material_indices(ctypes.byref(c), out.ctypes.data_as(ctypes.POINTER(c_uint32)),
                                              out.size))
        return out

    def tables(self) -> Tuple[np.ndarray, np.ndarray]:
        """(coefficients [n_materials, 4], axis table [4, nx+ny+nz]) as the engine uploads them."""
        from ._lib import check, lib

        c = self.to_c()
        extent = self.grid.extent
        span = extent.x + extent.y + extent.z
        coefficients = np.zeros(4 * len(self.materials), dtype=np.float32)
        axis = np.zeros(4 * span, dtype=np.float32)
        fp = ctypes.POINTER(c_float)
        check(lib().dp_scene_tables(ctypes.byref(c), coefficients.ctypes.data_as(fp), coefficients.size,
                                    axis.ctypes.data_as(fp), axis.size))
        return coefficients.reshape(-1, 4), axis.reshape(4, span)

    def injections(self, time: float) -> List[DpInjection]:
        """Source::injections for every source, in order (two each)."""
        from ._lib import check, lib

        c = self.to_c()
        out = (DpInjection * max(2 * len(self.sources), 1))()
        check(lib().dp_scene_injections(ctypes.byref(c), float(f32(time)), out, len(out)))
        return list(out)[: 2 * len(self.sources)]
