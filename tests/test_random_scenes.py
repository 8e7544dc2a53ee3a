"""Seeded random scenes: the engine's host layer against the oracle (no GPU).

Forty draws of grids, boundaries, materials, shapes and sources; for each,
the resolved grid, the rasterized material indices, the coefficient and axis
tables and the per-step injections must equal the oracle's bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import AXES

from random_scenes import random_case

SEEDS = range(40)


def _injection(inj):
    return (tuple(inj.origin), tuple(inj.extent), inj.component, inj.magnetic,
            np.float32(inj.value).view(np.uint32), tuple(inj.center),
            np.float32(inj.inverse_waist_squared).view(np.uint32))


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_host_tables_match_the_oracle(seed):
    case = random_case(seed)
    scene = case.scene
    extent, widths, dt = oracle.grid_resolve(scene.grid)
    assert scene.grid.extent.as_array() == extent, case.describe()
    engine_widths = np.concatenate([scene.grid.primary(a) for a in AXES])
    assert np.array_equal(engine_widths.view(np.uint32), widths.view(np.uint32)), case.describe()
    assert np.float32(scene.grid.time_step()).view(np.uint32) == np.float32(dt).view(np.uint32)

    assert np.array_equal(scene.material_indices(), oracle.material_indices(scene)), case.describe()

    coefficients, axis = scene.tables()
    assert np.array_equal(coefficients.view(np.uint32),
                          oracle.material_coefficients(scene).view(np.uint32)), case.describe()
    span = sum(extent)
    geometry = oracle.packed_geometry(scene.grid)
    absorber = oracle.absorber_packed(scene)
    for row, expected in ((0, geometry[:span]), (1, geometry[span:2 * span]), (2, absorber[:span]),
                          (3, absorber[span:])):
        assert np.array_equal(axis[row].view(np.uint32), expected.view(np.uint32)), f"{case.describe()} row {row}"

    dt32 = np.float32(scene.grid.time_step())
    for step in (0, 1, case.steps // 3, case.steps - 1, case.steps):
        t = float(np.float32(step) * dt32)
        engine = [_injection(i) for i in scene.injections(t)]
        reference = [_injection(i) for i in oracle.scene_injections(scene, t)]
        assert engine == reference, f"{case.describe()} step {step}"


def test_the_draws_cover_every_kind():
    cases = [random_case(seed) for seed in SEEDS]
    sources = {s.shape for c in cases for s in c.scene.sources}
    waveforms = {s.waveform.kind for c in cases for s in c.scene.sources}
    shapes = {s.kind for c in cases for s in c.scene.shapes}
    boundaries = {c.scene.boundary.kind for c in cases}
    graded = sum(len(set(c.scene.grid.primary(a).tolist())) > 1 for c in cases for a in AXES)
    moving = sum(any(v != 0.0 for v in s.velocity) for c in cases for s in c.scene.sources)
    ragged = sum(c.scene.grid.extent.x % 64 != 0 for c in cases)
    assert sources == {0, 1, 2} and waveforms == {0, 1, 2, 3} and shapes == {0, 1, 2}
    assert boundaries == {"pec", "absorbing"}
    assert graded > 0 and moving > 0 and ragged > 30


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scenes_survive_a_ron_round_trip(seed):
    # Every construct the draws use (graded grids, both boundaries, all
    # shapes, sources and waveforms) written as RON and read back equal.
    from diaphane_amd import Scene

    case = random_case(seed)
    assert Scene.from_ron(case.scene.to_ron()) == case.scene, case.describe()
