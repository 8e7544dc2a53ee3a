"""Synthetic scenes for the benchmark and the parity suite.

Every generator is deterministic: randomness comes from a SplitMix64 stream
seeded per workload (the reference itself uses no randomness). Sizes follow
the workload descriptions:

* ``free_space_80``  — 80³ uniform vacuum, Δt at 0.99 of the 3D Courant limit,
  a +x TFSF Gaussian-modulated packet (λ = 16 cells, envelope σ = 3 periods),
  default absorber, 500 steps.
* ``graded_scattering_256`` — ≈256³ cells graded along x around a 0.1·L band
  (ratio cap 1.15), 24 random dielectric spheres plus a PEC sphere, TFSF +x
  Gaussian pulse at λ = 20 fine cells, default absorber, 2000 steps.
* ``layered_768`` — 768³ uniform grid, 40 random z-slabs from a 15-entry
  lossless palette (u8 indices into 16 materials), a centred z-dipole driven by
  a differentiated Gaussian of σ = 60 Δt, 1000 steps with |E|² accumulated over
  the last 500.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np

from .scene import (Axis, Boundary, Extent, Grid, Material, Refinement, Scene, Shape, Source, Waveform,
                    SPEED_OF_LIGHT, f32)

MASK = (1 << 64) - 1


class SplitMix64:
    """Steele, Lea & Flood's SplitMix64 generator."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] inclusive."""
        return lo + self.next_u64() % (hi - lo + 1)

    def log_uniform_int(self, lo: int, hi: int) -> int:
        return int(round(math.exp(self.uniform(math.log(lo), math.log(hi)))))


@dataclass
class Workload:
    name: str
    scene: Scene
    steps: int
    accumulate_from: Optional[int] = None  # step at which |E|² accumulation starts
    readback_every: Optional[int] = None
    note: str = ""


def free_space_80(side: int = 80) -> Workload:
    courant = 0.57157676  # 0.99 × 0.57735026
    grid = Grid(Extent.cube(side), 1e-3, courant)
    scene = Scene(grid, Boundary.DEFAULT)
    frequency = float(SPEED_OF_LIGHT / f32(16e-3))  # λ = 16 cells
    cycles = 3.0 / 0.4246609  # FWHM cycles for σ = 3 periods
    delay = 9.0 / frequency   # 3σ
    scene.sources.append(Source.plane_wave(Axis.X, -0.3 * side * 1e-3, Axis.Z,
                                           Waveform.GaussianPulse(frequency, cycles, delay)))
    return Workload("free_space_80", scene, 500, readback_every=5,
                    note="TFSF +x Gaussian packet, λ=16 cells, σ=3 periods, S=0.99·limit")


def _graded_grid(target_nx: int, side: int, cell: float) -> Grid:
    best = None
    for base in range(max(side // 2, 8), side + 1):
        length = base * cell
        grid = Grid(Extent(base, side, side), cell, 0.5,
                    [Refinement.across(Axis.X, 0.0, 0.1 * length, cell / 4)], 1.15)
        nx = grid.extent.x
        if best is None or abs(nx - target_nx) < abs(best[0] - target_nx):
            best = (nx, grid)
        if nx >= target_nx:
            break
    return best[1]


def graded_scattering_256(side: int = 256, seed: int = 2) -> Workload:
    rng = SplitMix64(seed)
    cell = 1e-3
    grid = _graded_grid(side, side, cell)
    scene = Scene(grid, Boundary.DEFAULT)
    size = grid.size()
    L = min(size)
    for _ in range(24):
        centre = [rng.uniform(-0.4 * s, 0.4 * s) for s in size]
        radius = rng.uniform(0.02, 0.06) * L
        eps = rng.uniform(1.0, 4.0)
        sigma = rng.uniform(0.0, 0.05)
        m = scene.materials.push(Material(eps, 1.0, sigma, 0.0))
        scene.shapes.append(Shape.Sphere(centre, radius, m))
    pec = scene.materials.push(Material.PERFECT_CONDUCTOR)
    scene.shapes.append(Shape.Sphere([0.05 * size[0], 0.0, 0.0], 0.1 * L, pec))
    fine = cell / 4
    frequency = float(SPEED_OF_LIGHT / (f32(20.0) * f32(fine)))
    scene.sources.append(Source.plane_wave(Axis.X, -0.35 * size[0], Axis.Z,
                                           Waveform.gaussian_pulse(frequency, 3.0)))
    return Workload("graded_scattering_256", scene, 2000,
                    note="x graded (cap 1.15) around a 0.1·L band at Δ/4; TFSF on the graded axis "
                         "as configured (Scene::validate would refuse it; Simulation::new does not)")


def layered_768(side: int = 768, seed: int = 3, nz: Optional[int] = None) -> Workload:
    rng = SplitMix64(seed)
    nz = side if nz is None else nz
    cell = 1e-3
    grid = Grid(Extent(side, side, nz), cell)
    scene = Scene(grid, Boundary.DEFAULT)
    palette = [scene.materials.push(Material.dielectric(rng.uniform(1.0, 6.0))) for _ in range(15)]
    for _ in range(40):
        thickness = rng.integer(4, 30)
        start = rng.integer(0, max(side - thickness, 0))
        # slabs are drawn on the full 768-cell stack; a shorter sample keeps
        # the same metres-from-centre geometry
        offset = ((start + 0.5 * thickness) - 0.5 * side) * cell
        material = palette[rng.integer(0, len(palette) - 1)]
        scene.shapes.append(Shape.Slab(Axis.Z, offset, thickness * cell, material))
    dt = grid.time_step()
    sigma = 60.0 * dt
    scene.sources.append(Source.point([0.0, 0.0, 0.0], Axis.Z, Waveform.DifferentiatedGaussian(sigma, 4.0 * sigma)))
    return Workload("layered_768", scene, 1000, accumulate_from=500,
                    note="40 z-slabs, 15-entry lossless palette (u8 index), centred z-dipole, "
                         "differentiated Gaussian σ=60Δt, intensity over the last 500 steps")


WORKLOADS: Dict[str, Callable[..., Workload]] = {
    "free_space_80": free_space_80,
    "graded_scattering_256": graded_scattering_256,
    "layered_768": layered_768,
}
