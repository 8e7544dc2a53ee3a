"""Two steps per launch (DP_TILE_TWO_STEP) against the oracle, bit for bit.

The two-step kernel keeps the first step's H and E in registers and runs the
second step one z-plane behind. Its fields must equal two single steps'
exactly: on the presets with odd advance counts (which leave single steps
between the pairs), on every zero-sign pattern in an absorbing box (a loss
above 1 flips the sign of a zero field, so a zero produced by a different
expression shows up at once), on random fields, and with the intensity sum on.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Extent, Scene
from diaphane_amd.gpu import TILE_BUFFER, TILE_NONTEMPORAL, TILE_STACK, TILE_TWO_STEP, TILE_XCD_REMAP, Simulation

pytestmark = pytest.mark.gpu

SHAPES = [(6, 1, TILE_NONTEMPORAL | TILE_TWO_STEP), (8, 3, TILE_TWO_STEP), (10, 16, TILE_BUFFER | TILE_TWO_STEP),
          (8, 0, TILE_BUFFER | TILE_NONTEMPORAL | TILE_TWO_STEP),
          # XCD-grouped workgroups: runs of 8, and of 3 (a tail past the last whole group)
          (10, 4, TILE_BUFFER | TILE_TWO_STEP | TILE_XCD_REMAP), (6, 2, TILE_TWO_STEP | TILE_XCD_REMAP | (3 << 8)),
          # waves stacked along y: 4 waves of 6 rows (20 owned), 2 of 4 (4 owned), pointer addressing
          (6, 0, TILE_BUFFER | TILE_TWO_STEP | TILE_STACK | TILE_XCD_REMAP),
          (4, 3, TILE_BUFFER | TILE_TWO_STEP | TILE_STACK | (2 << 16)),
          (4, 5, TILE_TWO_STEP | TILE_STACK | (4 << 16) | TILE_XCD_REMAP | (3 << 8)),
          (6, 1, TILE_TWO_STEP | TILE_STACK | (2 << 16))]
SHAPE_IDS = [f"rows{r}-depth{d}-flags{f}" for r, d, f in SHAPES]


def differing(a: np.ndarray, b: np.ndarray) -> int:
    bits = np.uint64 if a.dtype == np.float64 else np.uint32
    return int(np.count_nonzero(a.view(bits) != b.view(bits)))


def assert_same(reference, device, label: str, intensity: bool = False) -> None:
    assert reference.step_count() == device.step_count()
    assert differing(reference.electric_all, device.read_electric()) == 0, f"{label}: E"
    assert differing(reference.magnetic_all, device.read_magnetic()) == 0, f"{label}: H"
    if intensity:
        assert reference.accumulated_steps() == device.accumulated_steps()
        assert differing(reference.intensity_f64(), device.read_intensity_f64()) == 0, f"{label}: intensity"


@pytest.mark.parametrize("rows,depth,flags", SHAPES, ids=SHAPE_IDS)
@pytest.mark.parametrize("chunks", [[7], [1, 2, 3, 2], [5, 2, 1, 4]])
def test_pairs_and_leftover_single_steps_match_the_oracle(context, rows, depth, flags, chunks):
    scene = Scene.photon(Extent(48, 32, 40))
    reference = oracle.Simulation(scene)
    device = Simulation(context, scene)
    device.set_tile(rows, depth, flags)
    for k, chunk in enumerate(chunks):
        if k == len(chunks) // 2:
            reference.set_accumulate_intensity(True)
            device.set_accumulate_intensity(True)
        reference.advance_by(chunk)
        device.advance_by(chunk)
    assert_same(reference, device, f"photon {chunks}", intensity=len(chunks) > 1)
    device.close()


def zero_patterns(cells: int):
    rng = np.random.default_rng(3)
    plus, minus = np.float32(0.0), np.float32(-0.0)
    full = lambda v: np.full(3 * cells, v, np.float32)  # noqa: E731

    def one(component, v):
        a = full(plus)
        a[component * cells:(component + 1) * cells] = v
        return a

    yield "E-0", full(minus), full(plus)
    yield "H-0", full(plus), full(minus)
    yield "both-0", full(minus), full(minus)
    yield ("mixed", np.where(rng.random(3 * cells) < 0.5, plus, minus).astype(np.float32),
           np.where(rng.random(3 * cells) < 0.5, plus, minus).astype(np.float32))
    for c in range(3):
        yield f"Hc{c}-0", full(plus), one(c, minus)
        yield f"Ec{c}-0", one(c, minus), full(plus)


@pytest.mark.parametrize("rows,depth,flags", SHAPES, ids=SHAPE_IDS)
def test_the_sign_of_zero_fields_survives_the_absorber(context, rows, depth, flags):
    # A sourceless absorbing box holding only signed zeros: every difference
    # between the pair's expressions and the single steps' would show here.
    scene = Scene.photon(Extent(48, 32, 32))
    scene.sources.clear()
    cells = 48 * 32 * 32
    for name, electric, magnetic in zero_patterns(cells):
        reference = oracle.Simulation(scene)
        reference.restore_fields(electric, magnetic, 1)
        reference.advance_by(4)
        device = Simulation(context, scene)
        device.set_tile(rows, depth, flags)
        device.write_fields(electric, magnetic, 1)
        device.advance_by(4)
        assert_same(reference, device, name)
        device.close()


@pytest.mark.parametrize("rows,depth,flags", SHAPES, ids=SHAPE_IDS)
def test_random_fields_in_a_graded_lossy_scene(context, rows, depth, flags):
    from random_scenes import random_case

    case = random_case(3)  # absorbing box, four materials and shapes, graded spacing
    scene = case.scene
    cells = scene.grid.extent.total()
    rng = np.random.default_rng(5)
    electric = rng.standard_normal(3 * cells).astype(np.float32)
    magnetic = rng.standard_normal(3 * cells).astype(np.float32)
    reference = oracle.Simulation(scene)
    reference.restore_fields(electric, magnetic, 9)
    reference.set_accumulate_intensity(True)
    reference.advance_by(6)
    device = Simulation(context, scene)
    device.set_tile(rows, depth, flags)
    device.write_fields(electric, magnetic, 9)
    device.set_accumulate_intensity(True)
    device.advance_by(6)
    assert_same(reference, device, case.describe(), intensity=True)
    device.close()


def test_an_odd_advance_leaves_one_single_step(context):
    device = Simulation(context, Scene.photon(Extent(48, 32, 32)))
    device.set_tile(8, 0, TILE_TWO_STEP)
    before = device.launch_count()
    device.advance_by(7)
    assert device.launch_count() - before == 4  # three pairs and one single step
    device.close()


def test_the_autotuner_considers_two_step_shapes_and_leaves_the_state_alone(context):
    scene = Scene.photon(Extent(96, 64, 64))
    device = Simulation(context, scene)
    device.advance_by(5)
    electric, magnetic = device.read_electric(), device.read_magnetic()
    tuned = device.autotune(1)
    assert tuned["rows"] in (4, 6, 8, 10) and tuned["ms_per_step"] > 0.0
    assert differing(electric, device.read_electric()) == 0 and differing(magnetic, device.read_magnetic()) == 0
    reference = oracle.Simulation(scene)
    reference.advance_by(12)
    device.advance_by(7)
    assert_same(reference, device, f"after autotune picked {tuned}")
    device.close()


@pytest.mark.parametrize("seed", range(100, 124))
def test_random_scenes_in_pairs_match_the_oracle(context, seed):
    # The seeded random scenes (graded grids, both boundaries, every shape,
    # source and waveform) advanced two steps per launch, split at an odd
    # step with the intensity sum on for the second half of every other draw.
    from random_scenes import random_case

    case = random_case(seed)
    rng = np.random.default_rng(20_000 + seed)
    rows = int(rng.choice([6, 8, 10]))
    depth = int(rng.choice([0, 1, 5, 16, 64]))
    flags = TILE_TWO_STEP | int(rng.choice([0, TILE_NONTEMPORAL, TILE_BUFFER, TILE_NONTEMPORAL | TILE_BUFFER]))
    intensity = bool(seed % 2)
    print(f"{case.describe()} rows {rows} depth {depth} flags {flags:#x}", flush=True)
    reference = oracle.Simulation(case.scene)
    device = Simulation(context, case.scene)
    device.set_tile(rows, depth, flags)
    first = case.steps // 2 | 1
    for sim in (reference, device):
        sim.advance_by(first)
        if intensity:
            sim.set_accumulate_intensity(True)
        sim.advance_by(case.steps - first)
    assert np.abs(reference.electric_all).max() > 0.0, "nothing was injected"
    assert_same(reference, device, case.describe(), intensity)
    device.close()


def test_large_grids_default_to_two_steps_per_launch(context):
    # At 32 Mi cells and above (with buffer addressing) a new sim starts two
    # steps per launch; its fields equal one step per launch bit for bit.
    scene = Scene.photon(Extent(320, 320, 328))  # 33.6M cells
    device = Simulation(context, scene)
    assert device.cells >= 32 << 20
    before = device.launch_count()
    device.advance_by(5)
    assert device.launch_count() - before == 3  # two pairs and a single step
    single = Simulation(context, scene)
    single.set_tile(4, 0, TILE_NONTEMPORAL)
    single.advance_by(5)
    assert differing(device.read_electric(), single.read_electric()) == 0
    assert differing(device.read_magnetic(), single.read_magnetic()) == 0
    for sim in (device, single):
        sim.set_accumulate_intensity(True)
        sim.advance_by(4)
    assert differing(device.read_intensity_f64(), single.read_intensity_f64()) == 0
    device.close()
    single.close()
    small = Simulation(context, Scene.photon(Extent(96, 64, 64)))
    before = small.launch_count()
    small.advance_by(5)
    assert small.launch_count() - before == 5  # one step per launch below the threshold
    small.close()
