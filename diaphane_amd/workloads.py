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
* ``batched_uneven_256`` — 256 independent solves, each axis drawn
  log-uniform in [48, 160] cells, half PEC cavities and half absorbing boxes,
  1–8 dielectric blocks (ε in [1, 4]) each, seeded with a random-phase
  superposition of 4 low-order discrete cavity modes; 1000 steps.
* ``served_stream`` — Poisson arrivals of solve requests, each axis drawn
  log-uniform in [32, 256] cells, step counts uniform in [200, 5000], scene
  types mixed evenly: PEC cavity ringdown, TFSF plane wave on a PEC sphere
  of random radius, and a matched-lossy half-space with σ in [0.01, 0.1].
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


@dataclass
class Instance:
    """One solve of a batched workload: its scene and initial fields."""

    scene: Scene
    modes: List[Tuple[Tuple[int, int, int], Tuple[float, float, float], float]]  # (m, amplitude, phase)

    def initial_fields(self) -> Tuple[np.ndarray, np.ndarray]:
        return cavity_mode_fields(self.scene.grid.extent.as_array(), self.modes)


def _mode_factors(n: int, m: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """1D factors of the discrete PEC mode m along an axis of n cells.

    sin(k·i) samples integer positions (zero on the wall at i = 0 and on the
    implicit wall at i = n), cos(k·(i+½)) the half positions; the forward
    differences are those of the H update (the value past the end is 0).
    """
    k = math.pi * m / n
    i = np.arange(n, dtype=np.float64)
    s = np.sin(k * i)
    c = np.cos(k * (i + 0.5))
    ds = np.append(s[1:], 0.0) - s
    dc = np.append(c[1:], 0.0) - c
    return s, c, ds, dc


def cavity_mode_fields(extent: Tuple[int, int, int], modes) -> Tuple[np.ndarray, np.ndarray]:
    """E and H (component-major 3N f32) of a superposition of discrete modes.

    Mode (mx, my, mz) has E_a = A_a·cos(k_a(i_a+½))·Π_{b≠a} sin(k_b·i_b) with
    k_a = π·m_a/n_a; A is projected orthogonal to (sin(k_a/2))_a, which makes
    the discrete divergence zero. H is the forward-difference curl of E scaled
    by 1/|K̃|, K̃_a = 2·sin(k_a/2). A mode with phase φ contributes cos φ·E and
    sin φ·H.
    """
    nx, ny, nz = extent
    n = (nx, ny, nz)
    terms = {("e", a): [] for a in range(3)}
    terms.update({("h", a): [] for a in range(3)})
    for m, amplitude, phase in modes:
        f = [_mode_factors(n[a], m[a]) for a in range(3)]
        s_half = np.array([math.sin(math.pi * m[a] / n[a] / 2) for a in range(3)])
        amp = np.array(amplitude, np.float64)
        amp = amp - s_half * (amp @ s_half) / (s_half @ s_half)
        norm = 2.0 * math.sqrt(float(s_half @ s_half))
        ce, sh = math.cos(phase), math.sin(phase)
        for a in range(3):
            # E_a: cos along a, sin along the others
            terms["e", a].append((ce * amp[a], [f[t][1] if t == a else f[t][0] for t in range(3)]))
            # H_a = d_b E_c − d_c E_b (forward differences), b = a+1, c = a+2
            b, c = (a + 1) % 3, (a + 2) % 3
            fac_c = [f[t][1] if t == c else f[t][0] for t in range(3)]
            fac_c[b] = f[b][2]
            fac_b = [f[t][1] if t == b else f[t][0] for t in range(3)]
            fac_b[c] = f[c][2]
            terms["h", a].append((sh / norm * amp[c], fac_c))
            terms["h", a].append((-sh / norm * amp[b], fac_b))

    def assemble(field: str) -> np.ndarray:
        out = np.empty((3, nz * ny, nx), np.float32)
        for a in range(3):
            # Σ_t w_t·fz_t⊗fy_t⊗fx_t as one (nz·ny × T) @ (T × nx) product
            zy = np.stack([w * np.outer(fz, fy).reshape(-1) for w, (fx, fy, fz) in terms[field, a]], axis=1)
            xs = np.stack([fx for _, (fx, _, _) in terms[field, a]], axis=0)
            out[a] = zy @ xs
        return out.reshape(-1)

    return assemble("e"), assemble("h")


def batched_uneven(count: int = 256, seed: int = 4, lo: int = 48, hi: int = 160,
                   cell: float = 1e-3) -> List[Instance]:
    rng = SplitMix64(seed)
    out: List[Instance] = []
    for _ in range(count):
        extent = Extent(rng.log_uniform_int(lo, hi), rng.log_uniform_int(lo, hi), rng.log_uniform_int(lo, hi))
        grid = Grid(extent, cell)
        boundary = Boundary.PEC if rng.integer(0, 1) == 0 else Boundary.DEFAULT
        scene = Scene(grid, boundary)
        size = grid.size()
        for _ in range(rng.integer(1, 8)):
            m = scene.materials.push(Material.dielectric(rng.uniform(1.0, 4.0)))
            dims = [rng.uniform(0.1, 0.5) * s for s in size]
            centre = [rng.uniform(-0.5 * s + 0.5 * d, 0.5 * s - 0.5 * d) for s, d in zip(size, dims)]
            scene.shapes.append(Shape.Block(centre, dims, m))
        modes = []
        for _ in range(4):
            mode = (rng.integer(1, 2), rng.integer(1, 2), rng.integer(1, 2))
            amplitude = (rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1))
            modes.append((mode, amplitude, rng.uniform(0.0, 2.0 * math.pi)))
        out.append(Instance(scene, modes))
    return out


def batched_uneven_256(count: int = 256, seed: int = 4) -> List[Instance]:
    return batched_uneven(count, seed)


@dataclass
class Request:
    """One request of a served stream."""

    arrival_s: float
    kind: str
    scene: Scene
    steps: int


SERVED_KINDS = ("cavity_ringdown", "tfsf_sphere", "lossy_half_space")


def served_scene(kind: str, extent: Extent, rng: SplitMix64, cell: float = 1e-3) -> Scene:
    grid = Grid(extent, cell)
    size = grid.size()
    if kind == "cavity_ringdown":
        # a Ricker kick in a closed PEC box, then free ringing
        scene = Scene(grid, Boundary.PEC)
        at = [rng.uniform(-0.3, 0.3) * s for s in size]
        scene.sources.append(Source.point(at, Axis.Z, Waveform.ricker(grid.frequency_for_resolution(16.0))))
        return scene
    if kind == "tfsf_sphere":
        scene = Scene(grid, Boundary.DEFAULT)
        pec = scene.materials.push(Material.PERFECT_CONDUCTOR)
        radius = rng.uniform(0.1, 0.3) * min(size)
        scene.shapes.append(Shape.Sphere([0.0, 0.0, 0.0], radius, pec))
        offset = -0.5 * size[0] + (Boundary.DEFAULT.thickness + 2) * cell
        scene.sources.append(Source.plane_wave(Axis.X, offset, Axis.Z, Waveform.gaussian_pulse(
            grid.frequency_for_resolution(20.0), 3.0)))
        return scene
    if kind == "lossy_half_space":
        scene = Scene(grid, Boundary.DEFAULT)
        lossy = scene.materials.push(Material.matched_lossy(rng.uniform(1.0, 4.0), rng.uniform(0.01, 0.1)))
        scene.shapes.append(Shape.Block([0.0, 0.0, -0.25 * size[2]], [2.0 * size[0], 2.0 * size[1], 0.5 * size[2]],
                                        lossy))
        scene.sources.append(Source.point([0.0, 0.0, 0.25 * size[2]], Axis.Z,
                                          Waveform.ricker(grid.frequency_for_resolution(16.0))))
        return scene
    raise ValueError(f"unknown request kind {kind!r}")


def served_stream(count: int = 200, rate: float = 50.0, seed: int = 5, lo: int = 32, hi: int = 256,
                  steps_lo: int = 200, steps_hi: int = 5000) -> List[Request]:
    """``count`` requests with exponential inter-arrival gaps of mean 1/rate;
    kinds cycle so the mix is even."""
    rng = SplitMix64(seed)
    t = 0.0
    out: List[Request] = []
    for k in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        extent = Extent(rng.log_uniform_int(lo, hi), rng.log_uniform_int(lo, hi), rng.log_uniform_int(lo, hi))
        kind = SERVED_KINDS[k % len(SERVED_KINDS)]
        scene = served_scene(kind, extent, rng)
        out.append(Request(t, kind, scene, rng.integer(steps_lo, steps_hi)))
    return out


WORKLOADS: Dict[str, Callable[..., Workload]] = {
    "free_space_80": free_space_80,
    "graded_scattering_256": graded_scattering_256,
    "layered_768": layered_768,
}
