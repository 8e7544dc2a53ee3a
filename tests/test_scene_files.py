"""Scene files: the RON format and the curated scenes — tests/scenes.rs, ported.

The curated scenes are the reference's scenes/*.ron, kept as fixtures in
tests/golden/scenes/ (copied by tests/golden/make_scenes.py). Parsing is the
engine's (dp_scene_from_ron); the geometry and the short runs are checked with
the engine's rasterization and the CPU oracle.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Grid, Material, Refinement, Scene, Shape, Source, Waveform

SCENES = sorted((Path(__file__).resolve().parent / "golden" / "scenes").glob("*.ron"))


def test_the_fixture_holds_the_fifteen_curated_scenes():
    assert len(SCENES) == 15


@pytest.mark.parametrize("name,scene", [
    ("photon", Scene.photon(Extent.cube(48))),
    ("cavity", Scene.cavity(Extent.cube(48))),
    ("slab", Scene.slab(Extent.cube(48), 1.8)),
])
def test_every_preset_survives_a_round_trip(name, scene):
    # tests/scenes.rs:31-43
    assert Scene.from_ron(scene.to_ron()) == scene, f"{name} changed on the way through RON"


def test_an_infinite_conductivity_survives_a_round_trip():
    # tests/scenes.rs:45-67
    scene = Scene.sized([0.024] * 3, 1000.0)
    metal = scene.materials.push(Material.PERFECT_CONDUCTOR)
    scene.shapes.append(Shape.Sphere([0.0] * 3, 0.006, metal))
    parsed = Scene.from_ron(scene.to_ron())
    assert parsed == scene
    assert np.isinf(parsed.materials.get(metal).conductivity)
    coefficients, _ = parsed.tables()  # (eg, el, mg, ml) per material
    assert coefficients[metal][0] == 0.0 and coefficients[metal][1] == 1.0


@pytest.mark.parametrize("path", SCENES, ids=[p.stem for p in SCENES])
def test_the_curated_scenes_parse_and_are_resolvable(path):
    # tests/scenes.rs:69-83
    scene = Scene.load(path)
    scene.validate()
    assert scene.sources, f"{path.name}: a scene with no source does nothing"


@pytest.mark.parametrize("path", SCENES, ids=[p.stem for p in SCENES])
def test_the_curated_scenes_actually_paint_their_geometry(path):
    # tests/scenes.rs:85-111
    scene = Scene.load(path)
    if len(scene.materials) == 1:
        assert not scene.shapes, f"{path.name}: shapes but only vacuum in the table"
        return
    indices = scene.material_indices()
    assert np.array_equal(indices, oracle.material_indices(scene))
    counts = np.bincount(indices, minlength=len(scene.materials))
    for material in range(1, len(scene.materials)):
        assert counts[material] > 0, f"{path.name}: material {material} is declared but paints no cells"


@pytest.mark.parametrize("path", SCENES, ids=[p.stem for p in SCENES])
def test_the_curated_scenes_run(path):
    # tests/scenes.rs:113-132, on the oracle
    simulation = oracle.Simulation(Scene.load(path).with_resolution(400.0))
    simulation.advance_by(30)
    assert simulation.is_finite(), f"{path.name}: diverged within 30 steps"
    assert simulation.energy().total() > 0.0, f"{path.name}: the source deposited nothing"


def test_a_saved_scene_reloads_identically(tmp_path):
    # tests/scenes.rs:134-141
    scene = Scene.slab(Extent.cube(32), 2.0)
    path = tmp_path / "diaphane-scene-round-trip.ron"
    scene.save(path)
    assert Scene.load(path) == scene


def test_every_construct_survives_a_round_trip():
    # The constructs the presets do not reach: refinements, a PEC boundary,
    # blocks, moving points, plane waves and every waveform.
    grid = Grid(Extent(40, 30, 30), 1e-3, 0.5, [Refinement((0.002, 0.0, 0.0), (0.01, 0.0, 0.0), 2.5e-4)], 1.2)
    scene = Scene(grid, Boundary.PEC)
    glass = scene.materials.push(Material.refractive(1.5))
    lossy = scene.materials.push(Material.matched_lossy(2.0, 0.05))
    scene.shapes += [Shape.Block([0.001, -0.002, 0.0], [0.004, 0.006, 0.008], glass),
                     Shape.Sphere([-0.005, 0.0, 0.001], 0.003, lossy), Shape.Slab(Axis.Y, 0.004, 0.002, glass)]
    scene.sources += [
        Source.point_moving([-0.01, 0.0, 0.0], [1.2e8, 0.0, -3e7], Axis.Z, Waveform.Ricker(2e10, 7e-11)),
        Source.plane_wave(Axis.X, -0.012, Axis.Y, Waveform.ContinuousWave(1e10, 2.5)),
        Source.sheet(Axis.Z, 0.003, 0.01, Axis.X, Waveform.GaussianPulse(1.5e10, 3.0, 4e-10)).with_amplitude(0.7),
        Source.point([0.0, 0.0, 0.0], Axis.Y, Waveform.DifferentiatedGaussian(1e-11, 5e-11)),
    ]
    text = scene.to_ron()
    assert Scene.from_ron(text) == scene
    assert Scene.from_ron(Scene.from_ron(text).to_ron()) == scene


MINIMAL = """
Scene(grid: Grid(extent: Extent(x: 20, y: 20, z: 20), cell_size: 0.001, courant: 0.5),
      boundary: Pec, materials: MaterialTable(materials: [Material(relative_permittivity: 1.0,
      relative_permeability: 1.0, conductivity: 0.0, magnetic_loss: 0.0)]), shapes: [],
      sources: [Source(shape: Point(at: (0.0, 0.0, 0.0)), polarization: Z,
                       waveform: Ricker(peak_frequency: 2e10, delay: 7e-11), amplitude: 1.0)])
"""


def test_serde_defaults_and_optional_names():
    scene = Scene.from_ron(MINIMAL)
    assert scene.grid.max_ratio == np.float32(1.15) and scene.grid.refinements == []
    assert scene.sources[0].velocity == (0.0, 0.0, 0.0)
    # struct names are optional, fields come in any order, unknown fields are skipped
    nameless = MINIMAL.replace("Grid(extent", "(extent").replace("Source(shape", "(polarization: Z, "
                               "note: Something(with: [1, 2], nested: (3, 4)), shape").replace(
        "polarization: Z,\n                       waveform", "waveform")
    assert Scene.from_ron(nameless) == scene
    commented = "/* a /* nested */ comment */\n// a line comment\n" + MINIMAL.replace(
        "(0.0, 0.0, 0.0)", "[0.0, 0.0, 0.0], // a list is taken for a triple\n")
    assert Scene.from_ron(commented) == scene


@pytest.mark.parametrize("change,message", [
    (("courant: 0.5)", "courant: 0.5"), "expected"),
    (("cell_size: 0.001, ", ""), "missing field `cell_size`"),
    (("Extent(x: 20", "Extent(x: 1"), "at least 2 cells on every axis"),
    (("courant: 0.5", "courant: 0.6"), "courant must lie in"),
    (("cell_size: 0.001", "cell_size: -0.001"), "cell_size must be positive and finite"),
    (("polarization: Z", "polarization: W"), "unknown axis"),
    (("boundary: Pec", "boundary: Periodic"), "unknown boundary"),
    (("Extent(x: 20", "Extent(x: -20"), "unsigned integer"),
    (("shapes: []", "shapes: [], shapes: []"), "duplicate field `shapes`"),
])
def test_errors_name_the_problem_and_its_place(change, message):
    text = MINIMAL.replace(*change)
    with pytest.raises(ValueError, match=message) as error:
        Scene.from_ron(text)
    line, column = str(error.value).split(":")[:2]
    assert int(line) >= 1 and int(column) >= 1
