"""The 256³ graded scattering grid's widths, pinned by a committed fixture.

tests/golden/graded_256_widths.json holds the oracle's `grade()` widths for
the grid as f32 bit patterns (written by tests/golden/make_graded_widths.py).
The reference ships no vector for these bits, so the fixture guards the
restatement against drift, and the invariants of the reference's own
graded-mesh tests (tests/graded.rs:136-166: the growth cap, walls that do not
move, the refined band at its cell size, centres that land in their own cell)
are asserted on the fixture's numbers directly. The engine's host layer must
produce the same bits.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from diaphane_amd import AXES, workloads

FIXTURE = Path(__file__).resolve().parent / "golden" / "graded_256_widths.json"


@pytest.fixture(scope="module")
def fixture():
    data = json.loads(FIXTURE.read_text())
    widths = np.array([int(h, 16) for h in data["widths_f32_hex"]], dtype=np.uint32).view(np.float32)
    dt = np.array([int(data["dt_f32_hex"], 16)], dtype=np.uint32).view(np.float32)[0]
    return data, widths, dt


@pytest.fixture(scope="module")
def grid():
    return workloads.graded_scattering_256().scene.grid


def split(data, widths):
    nx, ny, nz = data["extent"]
    return widths[:nx], widths[nx:nx + ny], widths[nx + ny:]


def test_the_recipe_is_the_workloads(fixture, grid):
    data, _, _ = fixture
    spec = grid.spec()
    assert [int(v) for v in spec.extent] == data["recipe"]["base_extent"]
    assert np.float32(spec.cell_size) == np.float32(data["recipe"]["cell_size"])
    assert np.float32(spec.max_ratio) == np.float32(data["recipe"]["max_ratio"])


def test_the_oracle_reproduces_the_fixture(fixture, grid):
    data, widths, dt = fixture
    extent, resolved, resolved_dt = oracle.grid_resolve(grid)
    assert list(extent) == data["extent"]
    assert np.array_equal(resolved.view(np.uint32), widths.view(np.uint32))
    assert np.float32(resolved_dt).view(np.uint32) == dt.view(np.uint32)


def test_the_engine_reproduces_the_fixture(fixture, grid):
    data, widths, dt = fixture
    assert list(grid.extent.as_array()) == data["extent"]
    engine = np.concatenate([grid.primary(a) for a in AXES])
    assert np.array_equal(engine.view(np.uint32), widths.view(np.uint32))
    assert np.float32(grid.time_step()).view(np.uint32) == dt.view(np.uint32)


def test_the_growth_cap_holds(fixture):
    data, widths, _ = fixture
    for w in split(data, widths):
        w = w.astype(np.float64)
        ratio = np.maximum(w[1:] / w[:-1], w[:-1] / w[1:])
        assert ratio.max() <= 1.15 + 1e-5, f"neighbouring cells grew by {ratio.max()}"


def test_the_walls_do_not_move(fixture):
    data, widths, _ = fixture
    recipe = data["recipe"]
    for axis, w in enumerate(split(data, widths)):
        size = float(np.sum(w.astype(np.float64)))
        domain = recipe["base_extent"][axis] * recipe["cell_size"]
        assert abs(size - domain) < 1e-6, f"axis {axis} became {size} m"


def test_the_band_is_refined_and_the_rest_is_not(fixture):
    data, widths, _ = fixture
    recipe = data["recipe"]
    wx, wy, wz = (w.astype(np.float64) for w in split(data, widths))
    fine = recipe["refinement"]["cell_size"]
    coarse = recipe["cell_size"]
    # tests/graded.rs:156
    assert wx.min() < 1.2 * fine and wx.max() > 0.9 * coarse
    # Cell faces along x, from the low wall at −L/2. The march samples the
    # target at each cell's start and the final rescale (src/grid.rs:876-882)
    # moves faces by up to |1 − scale|·L/2 (about 0.07 mm here), so a transition
    # cell may sit just inside the band's edge; 0.5 mm in from each edge every
    # cell is the band's.
    faces = -0.5 * float(np.sum(wx)) + np.concatenate([[0.0], np.cumsum(wx)])
    centre, half = recipe["refinement"]["center"][0], 0.5 * recipe["refinement"]["size"][0]
    margin = 0.5e-3
    inside = (faces[:-1] >= centre - half + margin) & (faces[1:] <= centre + half - margin)
    assert inside.sum() >= int((2 * half - 2 * margin) / fine) - 2
    assert np.all(wx[inside] <= fine * (1 + 1e-5))
    assert np.all(wx <= coarse * (1 + 1e-5))
    # y and z carry no refinement: every cell is the base cell
    assert np.all(wy == np.float32(coarse)) and np.all(wz == np.float32(coarse))


@pytest.mark.parametrize("fraction", [0.0, 0.1, 0.5, 0.9, 1.0])
def test_every_cell_centre_lands_in_its_own_cell(fixture, grid, fraction):
    data, widths, _ = fixture
    cells = []
    for axis, w in enumerate(split(data, widths)):
        w = w.astype(np.float64)
        i = int((len(w) - 1) * fraction)
        low = -0.5 * float(np.sum(w)) + float(np.sum(w[:i]))
        cells.append((i, low + 0.5 * w[i]))
    index = tuple(i for i, _ in cells)
    assert oracle.cell_containing(grid, [c for _, c in cells]) == index
