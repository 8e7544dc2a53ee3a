"""The engine's C++ host layer against the CPU oracle, bit for bit (no GPU).

Grid resolution (grade, widths, Δt), material coefficients, gains and
absorber profiles, rasterized material index sets and per-step injections are
computed independently by the engine (diaphane_amd/csrc/scene.cpp) and by the
oracle (oracle/oracle.c); every value must agree exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import (AXES, Axis, Boundary, Extent, Grid, Material, Refinement, Scene, Shape, Source,
                          Waveform)
from diaphane_amd import workloads


def grids():
    return {
        "uniform": Grid.new(Extent(40, 60, 80), 1e-3),
        "courant099": Grid(Extent.cube(80), 1e-3, 0.57157676),
        "graded_xy": Grid.graded([48e-3, 40e-3, 36e-3], 1e3, [
            Refinement.across(Axis.X, 6e-3, 8e-3, 0.25e-3), Refinement.across(Axis.Y, -4e-3, 6e-3, 0.5e-3)]),
        "graded_band": Grid.graded([0.096, 0.040, 0.040], 1e3, [Refinement.across(Axis.X, 0.0, 8e-3, 0.25e-3)]),
        "graded_gentle": Grid.graded_at([0.096, 0.040, 0.040], 1e3,
                                        [Refinement.across(Axis.X, 0.0, 0.028, 0.25e-3)], 1.05),
        "graded_box": Grid.graded([0.032, 0.028, 0.024], 1e3, [
            Refinement([2e-3, -1e-3, 0.0], [6e-3, 4e-3, 3e-3], 0.3e-3)]),
        "graded_256": workloads.graded_scattering_256().scene.grid,
    }


@pytest.mark.parametrize("name", list(grids()))
def test_grid_resolution_matches_the_oracle(name):
    grid = grids()[name]
    extent, widths, dt = oracle.grid_resolve(grid)
    assert grid.extent.as_array() == extent
    engine_widths = np.concatenate([grid.primary(a) for a in AXES])
    assert np.array_equal(engine_widths.view(np.uint32), widths.view(np.uint32))
    assert np.float32(grid.time_step()) == np.float32(dt)


def scenes():
    s1 = Scene.empty(Extent(44, 40, 40), 1e-3)
    glass = s1.materials.push(Material.refractive(1.5))
    lossy = s1.materials.push(Material.matched_lossy(2.0, 0.05))
    metal = s1.materials.push(Material.PERFECT_CONDUCTOR)
    s1.shapes += [Shape.Slab(Axis.X, 6e-3, 8e-3, glass), Shape.Sphere([-8e-3, 0.0, 0.0], 5e-3, lossy),
                  Shape.Block([14e-3, 0.0, 0.0], [4e-3, 12e-3, 12e-3], metal)]
    s1.sources.append(Source.point([-10e-3, 0.0, 0.0], Axis.Z, Waveform.ricker(3e10)))
    s2 = Scene.on_grid(grids()["graded_xy"]).with_boundary(Boundary.absorbing(8, 1e-6))
    g = s2.materials.push(Material.refractive(1.5))
    s2.shapes.append(Shape.Slab(Axis.X, 6e-3, 2e-3, g))
    s2.shapes.append(Shape.Sphere([1e-3, -3e-3, 2e-3], 4e-3, s2.materials.push(Material(2.0, 1.5, 0.01, 3.0))))
    s2.sources.append(Source.sheet(Axis.Y, -8e-3, 6e-3, Axis.Z, Waveform.gaussian_pulse(2e10, 3.0)))
    s3 = Scene.photon(Extent(48, 36, 30))
    s3.sources.append(Source.point_moving([-6e-3, 2e-3, 0.0], [2.7e8, 0.0, 0.0], Axis.Z, Waveform.ricker(2e10)))
    s3.sources.append(Source.plane_wave(Axis.Y, -5e-3, Axis.X, Waveform.ContinuousWave(1e10, 2.0)))
    s3.sources.append(Source.plane_wave(Axis.Z, 3e-3, Axis.Y, Waveform.DifferentiatedGaussian(1e-11, 4e-11)))
    return {
        "materials": s1,
        "graded": s2,
        "sources": s3,
        "layered_small": workloads.layered_768(side=96, nz=96).scene,
        "graded_256_small": workloads.graded_scattering_256(side=64).scene,
        "free_space_80": workloads.free_space_80().scene,
    }


@pytest.mark.parametrize("name", list(scenes()))
def test_material_index_sets_match_the_oracle(name):
    scene = scenes()[name]
    assert np.array_equal(scene.material_indices(), oracle.material_indices(scene))


@pytest.mark.parametrize("name", list(scenes()))
def test_kernel_tables_match_the_oracle(name):
    scene = scenes()[name]
    coefficients, axis = scene.tables()
    assert np.array_equal(coefficients.view(np.uint32), oracle.material_coefficients(scene).view(np.uint32))
    extent = scene.grid.extent
    span = extent.x + extent.y + extent.z
    geometry = oracle.packed_geometry(scene.grid)
    absorber = oracle.absorber_packed(scene)
    assert np.array_equal(axis[0].view(np.uint32), geometry[:span].view(np.uint32))
    assert np.array_equal(axis[1].view(np.uint32), geometry[span:2 * span].view(np.uint32))
    assert np.array_equal(axis[2].view(np.uint32), absorber[:span].view(np.uint32))
    assert np.array_equal(axis[3].view(np.uint32), absorber[span:].view(np.uint32))


def _fields(inj):
    return (tuple(inj.origin), tuple(inj.extent), inj.component, inj.magnetic,
            np.float32(inj.value).view(np.uint32), tuple(inj.center),
            np.float32(inj.inverse_waist_squared).view(np.uint32))


@pytest.mark.parametrize("name", list(scenes()))
def test_injections_match_the_oracle(name):
    scene = scenes()[name]
    dt = scene.grid.time_step()
    for step in (0, 1, 7, 60, 160, 333, 1000):
        t = float(np.float32(step) * np.float32(dt))
        engine = [_fields(i) for i in scene.injections(t)]
        reference = [_fields(i) for i in oracle.scene_injections(scene, t)]
        assert engine == reference, f"step {step}"


def test_tfsf_index_sets_are_the_reference_rows():
    # src/source.rs:622-650: E row at the plane, H row one cell before it,
    # full cross-section, the cyclic sign.
    scene = Scene.empty(Extent(32, 40, 48), 1e-3)
    scene.sources.append(Source.plane_wave(Axis.X, 0.0, Axis.Z, Waveform.ContinuousWave(1e9, 1.0)))
    electric, magnetic = scene.injections(10.13e-9)
    assert (electric.magnetic, magnetic.magnetic) == (0, 1)
    assert (electric.component, magnetic.component) == (2, 1)
    assert electric.origin[0] == magnetic.origin[0] + 1 == 16
    assert tuple(electric.extent) == tuple(magnetic.extent) == (1, 40, 48)
    assert electric.value > 0.0 > magnetic.value


def test_validate_reports_the_reference_messages():
    scene = Scene.photon(Extent.cube(48))
    scene.sources = [Source.sheet(Axis.X, -0.01, 0.01, Axis.Z, Waveform.ricker(scene.grid.frequency_for_resolution(2.0)))]
    with pytest.raises(ValueError, match="cells per wavelength"):
        scene.validate()
    scene = Scene.empty(Extent.cube(24), 1e-3)
    scene.shapes.append(Shape.Block([0.0] * 3, [2e-3] * 3, 7))
    with pytest.raises(ValueError, match="material 7"):
        scene.validate()
    scene = Scene.sized([0.032] * 3, 1000.0)
    scene.shapes.append(Shape.Sphere([20.0, 20.0, 20.0], 5.0, scene.materials.push(Material.refractive(1.5))))
    with pytest.raises(ValueError, match="entirely outside"):
        scene.validate()
    scene = Scene.empty(Extent(32, 32, 32), 1e-3)
    scene.sources.append(Source.plane_wave(Axis.X, 0.0, Axis.X, Waveform.ricker(2e10)))
    with pytest.raises(ValueError, match="light is\\s+transverse|light is transverse"):
        scene.validate()
    scene = Scene.empty(Extent.cube(12), 1e-3)
    with pytest.raises(ValueError, match="meet in the middle"):
        scene.validate()
    Scene.empty(Extent.cube(24), 1e-3).validate()
