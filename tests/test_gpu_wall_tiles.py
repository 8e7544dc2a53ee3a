"""Two-step tiles at the walls, bit for bit against the oracle.

A stacked two-step tile runs every half-step on all its rows, so the first row
tile, whose rows 0 and 1 are the grid's, also owns them (their updates need
nothing below the low wall), and the last row tile, whose top row lies past
the high wall, owns the row below it. These tests start from random fields
(every cell non-zero, so a cell computed from a missing neighbour shows at
once) on grids whose height sits on each side of the points where the row
tile counts change, and whose width puts column seams and the walls at every
lane position that matters, in lossy and lossless boxes. They compare E, H
and the intensity sum with the oracle after an odd number of steps (pairs and
a leftover single step), for one-wave and stacked tiles, alone and batched.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Material, Scene, Shape, Source, Waveform
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import TILE_BUFFER, TILE_STACK, TILE_TWO_STEP, TILE_XCD_REMAP, Simulation

pytestmark = pytest.mark.gpu

# one wave per tile (6, 8, 10 rows); stacked: 4 waves of 4 rows (16 rows), 2 of 4 (8), 4 of 6 (24), 2 of 6 (12)
SHAPES = [(6, 3, TILE_TWO_STEP), (8, 0, TILE_TWO_STEP | TILE_BUFFER), (10, 5, TILE_TWO_STEP | TILE_BUFFER),
          (4, 4, TILE_TWO_STEP | TILE_STACK | TILE_BUFFER | (4 << 16) | TILE_XCD_REMAP),
          (4, 3, TILE_TWO_STEP | TILE_STACK | (2 << 16)),
          (6, 2, TILE_TWO_STEP | TILE_STACK | TILE_BUFFER | (4 << 16)),
          (6, 0, TILE_TWO_STEP | TILE_STACK | (2 << 16))]
SHAPE_IDS = [f"rows{r}-depth{d}-flags{f}" for r, d, f in SHAPES]

# widths on both sides of the column tile counts' steps (60 | 61, 120 | 121),
# heights on both sides of the stacked row tile counts' steps (16 rows: 15 | 16, 27 | 28;
# 8 rows: 7 | 8, 11 | 12; 12 rows: 11 | 12, 19 | 20; 24 rows: 23 | 24, 43 | 44)
EXTENTS = [(63, 15, 9), (64, 16, 8), (120, 27, 7), (121, 28, 9), (125, 12, 6), (61, 44, 7), (60, 43, 6),
           (5, 20, 5), (184, 11, 6), (185, 24, 5), (6, 7, 5), (66, 8, 5)]
# a thin absorbing layer, so that it fits the smallest of these grids
ABSORBING = Boundary.absorbing(2, 1e-4)


def wall_scene(extent, boundary) -> Scene:
    nx, ny, nz = extent
    scene = Scene.empty(Extent(nx, ny, nz), 1e-3)
    scene.boundary = boundary
    size = scene.grid.size()
    centre = lambda i, n, s: (i + 0.5) * 1e-3 - 0.5 * s  # noqa: E731
    frequency = scene.grid.frequency_for_resolution(12.0)
    lossy = scene.materials.push(Material.matched_lossy(2.5, 0.3))
    glass = scene.materials.push(Material.refractive(1.7))
    # a lossy block over the first column seam and the low walls, glass at the high ones
    scene.shapes.append(Shape.Block([centre(1, nx, size[0]), centre(1, ny, size[1]), 0.0],
                                    [4e-3, 4e-3, 4e-3], lossy))
    scene.shapes.append(Shape.Block([centre(nx - 2, nx, size[0]), centre(ny - 2, ny, size[1]), 0.0],
                                    [3e-3, 3e-3, 3e-3], glass))
    for x in sorted({0, 1, 58, 59, 60, 61, nx - 2, nx - 1}):
        for y in sorted({0, 1, ny - 2, ny - 1}):
            if 0 <= x < nx and 0 <= y < ny:
                at = [centre(x, nx, size[0]), centre(y, ny, size[1]), centre(nz // 2, nz, size[2])]
                scene.sources.append(Source.point(at, Axis((x + y) % 3), Waveform.ricker(frequency)))
    return scene


def random_state(cells: int, seed: int):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(3 * cells).astype(np.float32), rng.standard_normal(3 * cells).astype(np.float32))


def differing(a: np.ndarray, b: np.ndarray) -> int:
    bits = np.uint64 if a.dtype == np.float64 else np.uint32
    return int(np.count_nonzero(a.view(bits) != b.view(bits)))


def reference_run(scene, electric, magnetic, steps):
    reference = oracle.Simulation(scene)
    reference.restore_fields(electric, magnetic, 5)
    reference.set_accumulate_intensity(True)
    reference.advance_by(steps)
    return reference


@pytest.mark.parametrize("boundary", [ABSORBING, Boundary.PEC], ids=["absorbing", "pec"])
@pytest.mark.parametrize("extent", EXTENTS, ids=[f"{x}x{y}x{z}" for x, y, z in EXTENTS])
def test_wall_tiles_match_the_oracle(context, extent, boundary):
    scene = wall_scene(extent, boundary)
    cells = scene.grid.extent.total()
    electric, magnetic = random_state(cells, sum(extent))
    reference = reference_run(scene, electric, magnetic, 7)
    for rows, depth, flags in SHAPES:
        device = Simulation(context, scene)
        device.set_tile(rows, depth, flags)
        device.write_fields(electric, magnetic, 5)
        device.set_accumulate_intensity(True)
        device.advance_by(7)
        label = f"{extent} rows {rows} depth {depth} flags {flags}"
        assert differing(reference.electric_all, device.read_electric()) == 0, f"{label}: E"
        assert differing(reference.magnetic_all, device.read_magnetic()) == 0, f"{label}: H"
        assert differing(reference.intensity_f64(), device.read_intensity_f64()) == 0, f"{label}: intensity"
        device.close()


@pytest.mark.parametrize("rows,flags", [(6, TILE_TWO_STEP | TILE_BUFFER | TILE_XCD_REMAP | (4 << 8)),
                                        (8, TILE_TWO_STEP),
                                        (4, TILE_TWO_STEP | TILE_STACK | TILE_BUFFER | (4 << 16)),
                                        (6, TILE_TWO_STEP | TILE_STACK | (2 << 16))],
                         ids=["pair6", "pair8", "stack4x4", "stack2x6"])
def test_batch_members_at_the_walls_match_the_oracle(context, rows, flags):
    scenes = [wall_scene(extent, ABSORBING if k % 2 else Boundary.PEC) for k, extent in enumerate(EXTENTS)]
    states = [random_state(s.grid.extent.total(), 40 + k) for k, s in enumerate(scenes)]
    batch = Batch(context)
    sims = []
    for scene, (electric, magnetic) in zip(scenes, states):
        sim = Simulation(context, scene)
        sim.write_fields(electric, magnetic, 5)
        sim.set_accumulate_intensity(True)
        batch.add(sim)
        sims.append(sim)
    batch.set_tile(rows, 3, flags)
    batch.advance_by(9)
    batch.wait()
    for scene, (electric, magnetic), sim in zip(scenes, states, sims):
        reference = reference_run(scene, electric, magnetic, 9)
        label = f"{scene.grid.extent.as_array()} rows {rows} flags {flags}"
        assert differing(reference.electric_all, sim.read_electric()) == 0, f"{label}: E"
        assert differing(reference.magnetic_all, sim.read_magnetic()) == 0, f"{label}: H"
        assert differing(reference.intensity_f64(), sim.read_intensity_f64()) == 0, f"{label}: intensity"
    batch.close()
    for sim in sims:
        sim.close()
