"""The oracle's scene layer against the known answers of the reference's own
unit tests (tests/golden/known_answers.json: src/source.rs, src/boundary.rs
and src/material.rs tests), and the engine's host layer against the same
answers where it exposes them."""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Grid, Material, Scene, Source, Waveform

GOLDEN = json.loads((Path(__file__).parent / "golden" / "known_answers.json").read_text())


def _grid(spec) -> Grid:
    return Grid.new(Extent(*spec["extent"]), spec["cell_size"])


def _scene_with(source: Source, grid_spec) -> Scene:
    scene = Scene(_grid(grid_spec), Boundary.PEC)
    scene.sources.append(source)
    return scene


# ---------------------------------------------------------------------------
# Injections (src/source.rs tests)
# ---------------------------------------------------------------------------
INJ = GOLDEN["injections"]


@pytest.mark.parametrize("layer", ["oracle", "engine"])
def test_a_point_injection_covers_exactly_one_cell(layer):
    case = INJ["point"]
    scene = _scene_with(Source.point(case["at"], Axis(case["polarization"]), Waveform.ricker(1e9)), INJ["grid"])
    q = (oracle.scene_injections(scene, 0.0) if layer == "oracle" else scene.injections(0.0))[0]
    assert list(q.origin) == case["origin"]
    assert list(q.extent) == case["extent"]
    assert q.component == case["component"]
    assert oracle.injection_weight(scene, q, case["origin"]) == case["weight"]


def test_a_sheet_spans_the_plane_and_apodizes_in_metres():
    case = INJ["sheet"]
    scene = _scene_with(Source.sheet(Axis(case["axis"]), case["offset"], case["waist"], Axis(case["polarization"]),
                                     Waveform.ricker(1e9)), INJ["grid"])
    q = oracle.scene_injections(scene, 0.0)[0]
    assert list(q.origin) == case["origin"]
    assert list(q.extent) == case["extent"]
    # The taper is exp(−r²/w²) with r the transverse distance (metres) of the
    # cell centre from the sheet's axis; the normal does not take part.
    grid = scene.grid
    centres = [(np.arange(n) + 0.5) * 1e-3 - 0.5 * n * 1e-3 for n in INJ["grid"]["extent"]]
    waist = case["waist"]
    for y, z in ((20, 24), (25, 24), (19, 30), (0, 0), (39, 47)):
        r2 = centres[1][y] ** 2 + centres[2][z] ** 2
        expected = math.exp(-r2 / waist ** 2)
        got = oracle.injection_weight(scene, q, [8, y, z])
        assert abs(got - expected) < 2e-6 * max(1.0, expected), (y, z, got, expected)
    # on the axis of the sheet (the centre falls between cells): the weight
    # is largest next to it and never above 1
    weights = [oracle.injection_weight(scene, q, [8, y, 24]) for y in range(40)]
    assert max(weights) <= case["weight_on_axis"]
    assert int(np.argmax(weights)) in (19, 20)


def test_an_injection_value_follows_the_waveform():
    case = INJ["amplitude"]
    waveform = Waveform.ricker(1e9)
    scene = _scene_with(Source.point([0.0, 0.0, 0.0], Axis.X, waveform).with_amplitude(case["amplitude"]),
                        INJ["grid"])
    q = oracle.scene_injections(scene, case["time"])[0]
    expected = case["amplitude"] * oracle.waveform_evaluate(waveform, case["time"])
    assert abs(q.value - expected) < case["tolerance"]


def test_a_moving_point_tracks_its_velocity_and_goes_silent_outside():
    case = INJ["moving_point"]
    waveform = Waveform.ricker(1e9)
    scene = _scene_with(Source.point_moving([0.0, 0.0, 0.0], case["velocity"], Axis.Z, waveform), INJ["grid"])
    for time, origin in case["samples"]:
        assert list(oracle.scene_injections(scene, time)[0].origin) == origin
    delay = waveform.delay
    assert oracle.scene_injections(scene, delay)[0].value == 0.0
    sitter = _scene_with(Source.point([0.0, 0.0, 0.0], Axis.Z, waveform), INJ["grid"])
    assert abs(oracle.scene_injections(sitter, delay)[0].value) > 0.99


@pytest.mark.parametrize("layer", ["oracle", "engine"])
def test_a_plane_wave_makes_two_staggered_corrections(layer):
    case = INJ["plane_wave"]
    waveform = Waveform.ContinuousWave(case["frequency"], case["ramp_cycles"])
    scene = _scene_with(Source.plane_wave(Axis(case["axis"]), case["offset"], Axis(case["polarization"]), waveform),
                        INJ["grid"])
    pair = oracle.scene_injections(scene, case["time"]) if layer == "oracle" else scene.injections(case["time"])
    electric, magnetic = pair
    assert not electric.magnetic and magnetic.magnetic
    assert electric.component == case["electric_component"]
    assert magnetic.component == case["magnetic_component"]
    assert electric.origin[0] == magnetic.origin[0] + 1
    assert list(electric.extent) == case["extent"] and list(magnetic.extent) == case["extent"]
    assert np.sign(electric.value) == case["electric_sign"]
    assert np.sign(magnetic.value) == case["magnetic_sign"]


def test_a_sheet_past_the_far_wall_is_clamped():
    case = INJ["sheet_clamp"]
    scene = _scene_with(Source.sheet(Axis(case["axis"]), case["offset"], case["waist"], Axis(case["polarization"]),
                                     Waveform.ricker(1e9)), INJ["grid"])
    assert oracle.scene_injections(scene, 0.0)[0].origin[case["axis"]] == case["origin_along_axis"]


# ---------------------------------------------------------------------------
# Waveforms (src/source.rs tests) and the differentiated Gaussian
# ---------------------------------------------------------------------------
WAVE = GOLDEN["waveforms"]


def test_a_ricker_peaks_at_its_delay_and_has_zero_mean():
    case = WAVE["ricker"]
    waveform = Waveform.ricker(case["frequency"])
    delay = waveform.delay
    assert abs(oracle.waveform_evaluate(waveform, delay) - 1.0) < case["peak_tolerance"]
    assert abs(oracle.waveform_evaluate(waveform, 0.0)) < case["start_tolerance"]
    n = case["mean_samples"]
    span = np.float32(case["mean_span_delays"]) * np.float32(delay)
    total = np.float32(0.0)
    for i in range(n):
        total = np.float32(total + np.float32(oracle.waveform_evaluate(waveform, float(span * np.float32(i) / np.float32(n)))))
    assert abs(float(total) / n) < case["mean_tolerance"]


def test_a_gaussian_pulse_has_the_requested_width():
    case = WAVE["gaussian_pulse"]
    waveform = Waveform.gaussian_pulse(case["frequency"], case["cycles"])
    delay = waveform.delay
    period = 1.0 / case["frequency"]

    def envelope(periods_from_peak):
        centre = delay + periods_from_peak * period
        return max(abs(oracle.waveform_evaluate(waveform, centre + (i / 256.0 - 0.5) * period)) for i in range(256))

    assert abs(envelope(0.0) - 1.0) < case["peak_tolerance"]
    assert abs(envelope(0.5 * case["cycles"]) - case["half_width_value"]) < case["half_width_tolerance"]
    assert envelope(case["tail_cycles"]) < case["tail_bound"]
    assert abs(oracle.waveform_evaluate(waveform, 0.0)) < case["start_tolerance"]


def test_a_continuous_wave_ramps_from_rest():
    case = WAVE["continuous_wave"]
    waveform = Waveform.ContinuousWave(case["frequency"], case["ramp_cycles"])
    assert oracle.waveform_evaluate(waveform, -1e-12) == 0.0
    assert oracle.waveform_evaluate(waveform, 0.0) == 0.0
    assert abs(oracle.waveform_evaluate(waveform, case["quarter_period"])) < case["quarter_bound"]
    peak = max(abs(oracle.waveform_evaluate(waveform, case["late"] + i * case["late_step"]))
               for i in range(case["late_samples"]))
    assert abs(peak - 1.0) < case["late_peak_tolerance"]


def test_the_differentiated_gaussian_is_the_derivative_of_a_gaussian():
    # −u·exp(½(1−u²)), u = (t − delay)/σ: odd about the delay, zero there,
    # extrema ±1 at u = ∓1, and proportional to d/dt exp(−u²/2).
    sigma, delay = 60e-12, 240e-12
    waveform = Waveform.DifferentiatedGaussian(sigma, delay)
    value = lambda t: oracle.waveform_evaluate(waveform, t)  # noqa: E731
    assert value(delay) == 0.0
    assert abs(value(delay - sigma) - 1.0) < 1e-6
    assert abs(value(delay + sigma) + 1.0) < 1e-6
    ts = delay + sigma * np.linspace(-4, 4, 161)
    vs = np.array([value(float(t)) for t in ts])
    assert np.max(np.abs(vs + vs[::-1])) < 1e-5
    assert np.max(np.abs(vs)) <= 1.0 + 1e-6
    # numerical derivative of the Gaussian envelope, scaled to the same peak
    u = (ts - delay) / sigma
    derivative = -u * np.exp(-0.5 * u * u) * math.exp(0.5)
    assert np.max(np.abs(vs - derivative)) < 1e-5


# ---------------------------------------------------------------------------
# Absorber (src/boundary.rs tests)
# ---------------------------------------------------------------------------
ABS = GOLDEN["absorber"]


def _profiles(boundary):
    scene = Scene(_grid(ABS["grid"]), boundary)
    packed = oracle.absorber_packed(scene)
    n = ABS["grid"]["extent"]
    offsets = np.cumsum([0] + n)
    integer = [packed[offsets[a]:offsets[a + 1]] for a in range(3)]
    half = [packed[offsets[3] + offsets[a]:offsets[3] + offsets[a + 1]] for a in range(3)]
    return integer, half


def _loss(profiles, axis, coord, electric):
    integer, half = profiles
    pick = lambda t: (half if (t == axis) == electric else integer)[t]  # noqa: E731
    # the row split: (P1[y] + P2[z]) computed first, then added to P0[x]
    return np.float32(pick(0)[coord[0]]) + np.float32(np.float32(pick(1)[coord[1]]) + np.float32(pick(2)[coord[2]]))


def test_a_pec_boundary_has_no_loss():
    integer, half = _profiles(Boundary.PEC)
    assert all(not np.any(p) for p in integer + half)


def test_the_interior_is_untouched_and_corners_lose_most():
    profiles = _profiles(Boundary.DEFAULT)
    assert _loss(profiles, 0, ABS["interior"]["cell"], True) == 0.0
    assert _loss(profiles, 1, ABS["interior"]["cell"], False) == 0.0
    corner = _loss(profiles, 0, ABS["corner"], True)
    face = _loss(profiles, 0, ABS["face"], True)
    assert corner > face > 0.0


def test_the_loss_grows_monotonically_into_the_wall():
    case = ABS["monotone"]
    profiles = _profiles(Boundary.DEFAULT)
    previous = math.inf
    for i in range(case["cells"]):
        value = _loss(profiles, case["axis"], [i, *case["row"]], True)
        assert value <= previous
        previous = value
    assert _loss(profiles, case["axis"], [case["zero_at"], *case["row"]], True) == 0.0


def test_both_walls_absorb():
    case = ABS["walls"]
    profiles = _profiles(Boundary.DEFAULT)
    assert _loss(profiles, case["axis"], case["low"], False) > 0.0
    assert _loss(profiles, case["axis"], case["high"], False) > 0.0


def test_the_peak_matches_the_closed_form():
    case = ABS["peak"]
    integer, half = _profiles(Boundary.DEFAULT)
    peak = max(float(np.max(p)) for p in integer + half)
    order = case["grading_order"]
    expected = -(order + 1) * math.log(case["reflection"]) * case["courant"] / (4 * case["thickness"])
    assert abs(peak - expected) < case["tolerance"]


def test_electric_and_magnetic_losses_sample_different_positions():
    profiles = _profiles(Boundary.DEFAULT)
    cell = ABS["staggered"]["cell"]
    assert _loss(profiles, 0, cell, True) != _loss(profiles, 0, cell, False)


# ---------------------------------------------------------------------------
# Materials (src/material.rs tests)
# ---------------------------------------------------------------------------
MAT = GOLDEN["materials"]


def _coefficients(*materials):
    scene = Scene(_grid(MAT["grid"]), Boundary.PEC)
    for m in materials:
        scene.materials.push(m)
    return oracle.material_coefficients(scene)


def test_vacuum_is_lossless_with_equal_gains_and_courant_geometry():
    eg, el, mg, ml = _coefficients()[0]
    case = MAT["vacuum"]
    assert el == case["electric_loss"] and ml == case["magnetic_loss"]
    assert eg == mg == case["electric_gain"]
    grid = _grid(MAT["grid"])
    geometry = oracle.packed_geometry(grid)
    n = sum(MAT["grid"]["extent"])
    gains = geometry[: 2 * n]
    assert np.max(np.abs(gains - np.float32(grid.courant))) < 1e-6


def test_a_perfect_conductor_pins_the_field_to_zero():
    coefficients = _coefficients(Material.PERFECT_CONDUCTOR)
    eg, el = np.float32(coefficients[1][0]), np.float32(coefficients[1][1])
    for e, curl in MAT["pec_updates"]["cases"]:
        e, curl = np.float32(e), np.float32(curl)
        assert ((np.float32(1) - el) * e + eg * curl) / (np.float32(1) + el) == MAT["pec_updates"]["expected"]


def test_matched_loss_has_equal_rates():
    case = MAT["matched"]
    for permittivity in case["permittivities"]:
        _, el, _, ml = _coefficients(Material.matched_lossy(permittivity, case["conductivity"]))[1]
        assert abs(el - ml) / el < case["tolerance"]


def test_conductivity_raises_loss_without_touching_gain():
    case = MAT["conductivity"]
    lossless = _coefficients(Material.dielectric(case["permittivity"]))[1]
    lossy = _coefficients(Material.matched_lossy(case["permittivity"], case["conductivity"]))[1]
    assert lossless[0] == lossy[0]
    assert lossless[1] == 0.0 and lossy[1] > 0.0
