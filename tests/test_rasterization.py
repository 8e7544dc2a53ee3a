"""Rasterization, Scene::material_indices (src/scene.rs:181-206) with
Shape::contains (src/scene.rs:60-81): the reference's own tests
(src/scene.rs:510-629) on the oracle and on the engine's host layer, and a
brute-force check on grids whose every corner and centre is exact in binary,
where the answer is fixed by arithmetic alone: membership at the cell
centre, closed (<=) bounds, later shapes winning."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Extent, Grid, Material, MaterialTable, Scene, Shape

LAYERS = {"oracle": oracle.material_indices, "engine": lambda scene: scene.material_indices()}


@pytest.fixture(params=sorted(LAYERS))
def indices_of(request):
    return LAYERS[request.param]


def test_empty_scene_is_all_vacuum(indices_of):
    scene = Scene.empty(Extent.cube(24), 1e-3)
    indices = indices_of(scene)
    assert indices.size == 24 ** 3
    assert np.all(indices == MaterialTable.VACUUM)
    scene.validate()


def test_shapes_rasterize_and_later_ones_win(indices_of):
    scene = Scene.empty(Extent.cube(16), 1e-3)
    glass = scene.materials.push(Material.refractive(1.5))
    metal = scene.materials.push(Material.PERFECT_CONDUCTOR)
    scene.shapes.append(Shape.Block([0.0] * 3, [8e-3] * 3, glass))
    scene.shapes.append(Shape.Sphere([0.0] * 3, 2e-3, metal))
    indices = indices_of(scene)
    at = lambda c: indices[scene.grid.extent.index(c)]
    assert at([0, 0, 0]) == MaterialTable.VACUUM
    assert at([5, 5, 5]) == glass
    assert at([8, 8, 8]) == metal  # the sphere's centre, painted after the block
    assert at([11, 11, 11]) == glass


def test_slab_shape_spans_one_axis_only(indices_of):
    scene = Scene.empty(Extent.cube(12), 1e-3)
    glass = scene.materials.push(Material.refractive(2.0))
    scene.shapes.append(Shape.Slab(Axis.Y, -1.5e-3, 3e-3, glass))
    indices = indices_of(scene)
    at = lambda c: indices[scene.grid.extent.index(c)]
    assert at([0, 4, 11]) == glass
    assert at([11, 4, 0]) == glass
    assert at([5, 2, 5]) == MaterialTable.VACUUM
    assert at([5, 6, 5]) == MaterialTable.VACUUM


def test_geometry_occupies_the_same_volume_at_any_resolution(indices_of):
    scene = Scene.sized([0.032] * 3, 1000.0)
    glass = scene.materials.push(Material.refractive(1.5))
    metal = scene.materials.push(Material.PERFECT_CONDUCTOR)
    scene.shapes.append(Shape.Sphere([4e-3, -2e-3, 0.0], 6e-3, glass))
    scene.shapes.append(Shape.Block([-8e-3, 0.0, 0.0], [4e-3, 8e-3, 8e-3], metal))
    fine_scene = scene.with_resolution(2000.0)
    assert fine_scene.grid.extent == Extent.cube(64)
    for material in (glass, metal):
        coarse = float(np.mean(indices_of(scene) == material))
        fine = float(np.mean(indices_of(fine_scene) == material))
        assert coarse > 0.001, f"material {material} painted nothing"
        assert abs(fine - coarse) / coarse < 0.03, f"material {material}: {coarse:.4f} coarse vs {fine:.4f} fine"


# ---------------------------------------------------------------------------
# brute force on dyadic grids
# ---------------------------------------------------------------------------
CELL = 1.0 / 1024.0  # metres; every corner, centre and length below is exact in f32


def dyadic_scene(extent) -> Scene:
    scene = Scene.empty(Extent(*extent), CELL)
    a = scene.materials.push(Material.dielectric(2.0))
    b = scene.materials.push(Material.dielectric(3.0))
    c = scene.materials.push(Material.PERFECT_CONDUCTOR)
    # Each bound passes exactly through a row of cell centres, so a strict
    # `<` would lose that row; a block and a sphere overlap a slab that is
    # painted between them.
    scene.shapes.append(Shape.Block([2.0 * CELL, -1.0 * CELL, 0.5 * CELL], [9.0 * CELL, 5.0 * CELL, 11.0 * CELL], a))
    scene.shapes.append(Shape.Slab(Axis.Z, -2.5 * CELL, 3.0 * CELL, b))
    scene.shapes.append(Shape.Sphere([-3.5 * CELL, 2.5 * CELL, -1.5 * CELL], 5.0 * CELL, c))
    scene.shapes.append(Shape.Slab(Axis.X, 6.5 * CELL, 1.0 * CELL, a))
    return scene


def brute_force(scene: Scene) -> np.ndarray:
    nx, ny, nz = scene.grid.extent.as_array()
    # cell centres (i + ½)·Δ − n·Δ/2, exact in binary
    cx = (np.arange(nx) + 0.5 - nx / 2.0) * CELL
    cy = (np.arange(ny) + 0.5 - ny / 2.0) * CELL
    cz = (np.arange(nz) + 0.5 - nz / 2.0) * CELL
    z, y, x = np.meshgrid(cz, cy, cx, indexing="ij")
    p = (x, y, z)
    out = np.zeros((nz, ny, nx), dtype=np.uint32)
    for shape in scene.shapes:
        if shape.kind == 0:
            inside = np.ones_like(out, dtype=bool)
            for a in range(3):
                inside &= np.abs(p[a] - shape.center[a]) <= 0.5 * shape.size[a]
        elif shape.kind == 1:
            d2 = sum((p[a] - shape.center[a]) ** 2 for a in range(3))
            inside = d2 <= shape.radius ** 2
        else:
            inside = np.abs(p[int(shape.axis)] - shape.offset) <= 0.5 * shape.thickness
        out[inside] = shape.material
    return out.ravel()


@pytest.mark.parametrize("extent", [(24, 20, 22), (17, 31, 12)])
def test_membership_is_closed_at_the_cell_centre_and_later_shapes_win(indices_of, extent):
    scene = dyadic_scene(extent)
    expected = brute_force(scene)
    # the exact-boundary rows exist: a strict bound would paint fewer cells
    for shape_index in (0, 1, 3):
        shape = scene.shapes[shape_index]
        assert shape.kind in (0, 2)
    got = indices_of(scene)
    assert np.array_equal(got, expected), f"{np.count_nonzero(got != expected)} cells differ"
    counts = np.bincount(expected, minlength=4)
    assert all(counts[m] > 0 for m in range(4))


def test_a_strict_bound_would_be_caught():
    # The same brute force with `<` in place of `<=` paints fewer cells; the
    # layers agree with the closed form, so a `<` in either would fail above.
    scene = dyadic_scene((24, 20, 22))
    closed = brute_force(scene)
    nx, ny, nz = scene.grid.extent.as_array()
    cx = (np.arange(nx) + 0.5 - nx / 2.0) * CELL
    block = scene.shapes[0]
    on_bound = np.isclose(np.abs(cx - block.center[0]), 0.5 * block.size[0], rtol=0, atol=0)
    assert on_bound.any(), "the block's x bound should pass through cell centres"
    assert np.array_equal(oracle.material_indices(scene), closed)
