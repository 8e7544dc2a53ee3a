"""Seeded random scenes for parity tests.

Each draw combines a grid (uniform, or graded with one refinement), a
boundary (perfect conductor or an absorbing layer of random depth and target
reflection), up to four materials (dielectric, magnetic, lossy, matched
lossy, perfect conductor), up to five shapes (blocks, spheres, slabs) and one
to three sources (point, moving point, Gaussian sheet, plane wave) driven by
each waveform kind. Extents are ragged on purpose: x is seldom a multiple of
the kernel's 64-cell tile and the grids are small enough for the oracle to
finish in well under a second. A draw that the scene layer rejects
(too few cells per wavelength, a layer that does not fit) is drawn again.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from diaphane_amd import Axis, Boundary, Extent, Grid, Material, Refinement, Scene, Shape, Source, Waveform


@dataclass
class RandomCase:
    seed: int
    scene: Scene
    steps: int

    def describe(self) -> str:
        e = self.scene.grid.extent
        kinds = "+".join(str(s.shape) for s in self.scene.sources)
        return (f"seed {self.seed}: {e.x}x{e.y}x{e.z} {self.scene.boundary.kind}"
                f"({self.scene.boundary.thickness}) materials {len(self.scene.materials)} "
                f"shapes {len(self.scene.shapes)} sources {kinds} steps {self.steps}")


def _material(rng) -> Material:
    kind = rng.integers(5)
    if kind == 0:
        return Material.dielectric(float(rng.uniform(1.2, 6.0)))
    if kind == 1:
        return Material(float(rng.uniform(1.0, 4.0)), float(rng.uniform(1.0, 2.5)))
    if kind == 2:
        return Material(float(rng.uniform(1.0, 3.0)), 1.0, float(rng.uniform(0.0, 0.2)),
                        float(rng.uniform(0.0, 5.0)))
    if kind == 3:
        return Material.matched_lossy(float(rng.uniform(1.0, 3.0)), float(rng.uniform(0.001, 0.1)))
    return Material.PERFECT_CONDUCTOR


def _point_in(rng, size, fraction=0.35):
    return [float(rng.uniform(-fraction, fraction) * s) for s in size]


def _shape(rng, size, material: int) -> Shape:
    kind = rng.integers(3)
    if kind == 0:
        return Shape.Block(_point_in(rng, size), [float(rng.uniform(0.08, 0.4) * s) for s in size], material)
    if kind == 1:
        return Shape.Sphere(_point_in(rng, size), float(rng.uniform(0.05, 0.25) * min(size)), material)
    axis = Axis(int(rng.integers(3)))
    return Shape.Slab(axis, float(rng.uniform(-0.3, 0.3) * size[int(axis)]),
                      float(rng.uniform(0.04, 0.2) * size[int(axis)]), material)


def _waveform(rng, grid: Grid) -> Waveform:
    frequency = grid.frequency_for_resolution(float(rng.uniform(14.0, 30.0)))
    kind = rng.integers(4)
    if kind == 0:
        return Waveform.ricker(frequency)
    if kind == 1:
        return Waveform.gaussian_pulse(frequency, float(rng.uniform(1.0, 3.0)))
    if kind == 2:
        return Waveform.ContinuousWave(frequency, float(rng.uniform(0.5, 3.0)))
    width = 1.0 / (2.0 * np.pi * frequency)
    return Waveform.DifferentiatedGaussian(width, float(rng.uniform(3.0, 5.0)) * width)


def _source(rng, grid: Grid) -> Source:
    size = grid.size()
    waveform = _waveform(rng, grid)
    polarization = Axis(int(rng.integers(3)))
    kind = rng.integers(4)
    if kind == 0:
        source = Source.point(_point_in(rng, size, 0.3), polarization, waveform)
    elif kind == 1:
        velocity = [float(rng.uniform(-0.5, 0.5) * 3e8) for _ in range(3)]
        source = Source.point_moving(_point_in(rng, size, 0.2), velocity, polarization, waveform)
    else:
        axis = Axis(int(rng.integers(3)))
        if polarization == axis:
            polarization = Axis((int(axis) + 1 + int(rng.integers(2))) % 3)
        offset = float(rng.uniform(-0.25, 0.25) * size[int(axis)])
        if kind == 2:
            waist = float(rng.uniform(0.15, 0.4) * min(size))
            source = Source.sheet(axis, offset, waist, polarization, waveform)
        else:
            source = Source.plane_wave(axis, offset, polarization, waveform)
    return source.with_amplitude(float(rng.uniform(0.5, 2.0)))


def _grid(rng) -> Grid:
    extent = Extent(int(rng.integers(24, 80)), int(rng.integers(24, 56)), int(rng.integers(24, 56)))
    if rng.random() < 0.75:
        return Grid.new(extent, 1e-3)
    size = [1e-3 * n for n in (extent.x, extent.y, extent.z)]
    axis = Axis(int(rng.integers(3)))
    half = size[int(axis)] / 2
    start = float(rng.uniform(-0.6, 0.0) * half)
    width = float(rng.uniform(0.1, 0.4) * half)
    return Grid.graded(size, 1e3, [Refinement.across(axis, start, start + width,
                                                     float(rng.choice([0.25e-3, 0.5e-3])))])


def _draw(rng, seed: int) -> RandomCase:
    grid = _grid(rng)
    if rng.random() < 0.4:
        boundary = Boundary.PEC
    else:
        boundary = Boundary.absorbing(int(rng.integers(4, 11)), float(10.0 ** rng.uniform(-7, -3)))
    scene = Scene(grid, boundary)
    size = grid.size()
    for _ in range(int(rng.integers(0, 5))):
        scene.materials.push(_material(rng))
    for _ in range(int(rng.integers(0, 6))):
        if len(scene.materials) > 1:
            scene.shapes.append(_shape(rng, size, int(rng.integers(1, len(scene.materials)))))
    for _ in range(int(rng.integers(1, 4))):
        scene.sources.append(_source(rng, grid))
    return RandomCase(seed, scene, int(rng.integers(40, 160)))


def random_case(seed: int, attempts: int = 64) -> RandomCase:
    """The first draw from `seed`'s stream that the scene layer accepts."""
    rng = np.random.default_rng(seed)
    last = None
    for _ in range(attempts):
        case = _draw(rng, seed)
        try:
            case.scene.validate()
        except ValueError as error:
            last = error
            continue
        return case
    raise RuntimeError(f"seed {seed}: no valid draw in {attempts} attempts (last: {last})")
