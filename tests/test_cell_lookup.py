"""The renderer's world-to-cell table, Grid::cell_lookup
(src/grid.rs:635-646), and the sample count gpu::Simulation::new picks for
it (src/gpu.rs:225-227): the oracle pinned by closed forms, and the engine's
host layer bit-equal to the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Extent, Grid, Refinement

UNIFORM = [Grid.new(Extent(16, 24, 40), 1.0 / 1024.0), Grid.new(Extent(80, 80, 80), 1e-3),
           Grid.new(Extent(37, 53, 11), 2.5e-3)]


def graded_grids():
    return [
        Grid.graded([0.064, 0.032, 0.048], 1000.0, [Refinement.across(Axis.X, 0.0, 0.0064, 0.25e-3)]),
        Grid.graded([0.05, 0.05, 0.05], 1000.0,
                    [Refinement([0.01, -0.005, 0.0], [0.01, 0.008, 0.012], 0.3e-3)]),
    ]


@pytest.mark.parametrize("grid", UNIFORM + graded_grids())
def test_the_engine_table_is_the_oracle_table(grid):
    samples = grid.lookup_samples()
    a = grid.cell_lookup(samples)
    b = oracle.cell_lookup(grid, samples)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("grid", UNIFORM)
def test_a_uniform_axis_maps_position_to_cell_linearly(grid):
    # On a uniform axis the coordinate of the position (f − ½)·L is f·n: the
    # samples run from exactly 0 to exactly n in equal steps.
    samples = 300
    table = oracle.cell_lookup(grid, samples).reshape(3, samples)
    fraction = np.arange(samples, dtype=np.float64) / (samples - 1)
    for axis, n in enumerate(grid.extent.as_array()):
        row = table[axis].astype(np.float64)
        assert row[0] == 0.0
        if grid.base_cell_size == 1.0 / 1024.0:  # dyadic widths: every corner exact
            assert row[-1] == float(n)
        else:
            assert abs(row[-1] - n) < 1e-6 * n
        assert np.all(np.diff(row) > 0.0)
        assert np.max(np.abs(row - fraction * n)) < 2e-5 * n


@pytest.mark.parametrize("grid", graded_grids())
def test_a_graded_axis_inverts_the_cell_positions(grid):
    # Spacing::position is the inverse map; the round trip lands on the
    # sampled positions. Refined cells take proportionally more of the
    # coordinate range, so the table is steepest there.
    samples = grid.lookup_samples()
    table = oracle.cell_lookup(grid, samples).reshape(3, samples)
    for axis, n in enumerate(grid.extent.as_array()):
        length = grid.length(Axis(axis))
        row = table[axis]
        assert row[0] == 0.0 and abs(float(row[-1]) - n) < 1e-4 * n
        assert np.all(np.diff(row.astype(np.float64)) > 0.0)
        for s in range(0, samples, max(1, samples // 97)):
            target = (s / (samples - 1) - 0.5) * length
            back = oracle.position(grid, axis, float(row[s]))
            assert abs(back - target) < 2e-6 * length, (axis, s, back, target)
        widths = grid.primary(Axis(axis)).astype(np.float64)
        if widths.min() < 0.9 * widths.max():
            # cells per metre where the axis is finest vs coarsest
            corners = np.concatenate([[0.0], np.cumsum(widths)])
            slope = np.diff(row.astype(np.float64)) / (length / (samples - 1))
            fine = np.argmin(widths)
            at = int((corners[fine] + 0.5 * widths[fine]) / length * (samples - 1))
            assert slope[min(at, slope.size - 1)] > 0.95 / widths.min() * 0.99


@pytest.mark.parametrize("extent,expected", [((10, 10, 10), 256), ((64, 8, 8), 256), ((65, 8, 8), 260),
                                              ((100, 300, 20), 1200), ((2048, 2, 2), 8192),
                                              ((4000, 4, 4), 8192)])
def test_the_sample_count_is_four_per_cell_within_its_clamp(extent, expected):
    assert Grid.new(Extent(*extent), 1e-3).lookup_samples() == expected


def test_the_table_refuses_fewer_than_two_samples():
    from diaphane_amd._lib import DiaphaneError

    with pytest.raises(DiaphaneError):
        UNIFORM[0].cell_lookup(1)
    with pytest.raises(oracle.OracleError):
        oracle.cell_lookup(UNIFORM[0], 1)
