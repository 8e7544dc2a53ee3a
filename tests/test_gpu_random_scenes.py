"""Seeded random scenes on the device against the oracle, bit for bit.

Each draw (grid, boundary, materials, shapes, sources: see random_scenes.py)
runs through a kernel path and a tiling chosen from the same seed, with the
|E|² accumulator switched on halfway for every other draw; the fields and the
f64 intensity must equal the oracle's exactly. A second test advances
sixteen draws together in one batch, each to its own step count; a third
runs 300 more draws through the fused forms in chunks of random lengths, and
a fourth batches 24 draws at each batch tiling in chunks of random lengths.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import KERNEL_FUSED, KERNEL_FUSED_OVERLAP, KERNEL_TWO_PASS, TILE_BUFFER, TILE_NONTEMPORAL, \
    TILE_STACK, TILE_TWO_STEP, TILE_XCD_REMAP, Simulation

from random_scenes import random_case

pytestmark = pytest.mark.gpu

SOLO_SEEDS = range(100, 124)
BATCH_SEEDS = range(200, 216)


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def run_plan(seed: int):
    """(kernel, rows, depth, flags, accumulate) drawn from the seed."""
    rng = np.random.default_rng(10_000 + seed)
    kernel = (KERNEL_FUSED, KERNEL_FUSED, KERNEL_TWO_PASS, KERNEL_FUSED_OVERLAP)[int(rng.integers(4))]
    rows = int(rng.choice([4, 6, 8, 10, 12]))
    depth = int(rng.choice([0, 5, 16, 33, 64]))
    flags = int(rng.choice([0, TILE_NONTEMPORAL, TILE_BUFFER, TILE_NONTEMPORAL | TILE_BUFFER]))
    if rng.random() < 0.3:
        flags |= TILE_XCD_REMAP | (int(rng.choice([0, 2, 4])) << 8)
    if rng.random() < 0.5:
        flags |= TILE_TWO_STEP  # two steps per launch (the fused kernel's option)
    if flags & TILE_TWO_STEP and rng.random() < 0.5:
        flags |= TILE_STACK | (int(rng.choice([0, 2, 4])) << 16)  # waves stacked along y
    return kernel, rows, depth, flags, bool(seed % 2)


def assert_matches_oracle(case, reference, device, intensity: bool):
    label = case.describe()
    assert reference.step_count() == device.step_count() == case.steps
    electric, magnetic = device.read_electric(), device.read_magnetic()
    assert reference.is_finite(), f"{label}: the oracle diverged"
    assert np.abs(reference.electric_all).max() > 0.0, f"{label}: nothing was injected"
    for name, ref_all, cand in (("E", reference.electric_all, electric), ("H", reference.magnetic_all, magnetic)):
        diff = np.flatnonzero(ref_all.view(np.uint32) != cand.view(np.uint32))
        assert diff.size == 0, (f"{label} {name}: {diff.size} cells differ; first at {diff[0]}: "
                                f"oracle {ref_all[diff[0]]!r} vs device {cand[diff[0]]!r}")
    if intensity:
        assert reference.accumulated_steps() == device.accumulated_steps()
        assert np.array_equal(reference.intensity_f64(), device.read_intensity_f64()), f"{label}: intensity"


@pytest.mark.parametrize("seed", SOLO_SEEDS)
def test_a_random_scene_matches_the_oracle(context, seed):
    case = random_case(seed)
    kernel, rows, depth, flags, intensity = run_plan(seed)
    print(f"{case.describe()} kernel {kernel} rows {rows} depth {depth} flags {flags:#x}", flush=True)
    reference = oracle.Simulation(case.scene)
    device = Simulation(context, case.scene)
    device.set_kernel(kernel)
    device.set_tile(rows, depth, flags)
    half = case.steps // 2
    for sim in (reference, device):
        sim.advance_by(half)
        if intensity:
            sim.set_accumulate_intensity(True)
        sim.advance_by(case.steps - half)
    assert_matches_oracle(case, reference, device, intensity)
    device.close()


def test_random_scenes_batched_to_their_own_step_counts(context):
    cases = [random_case(seed) for seed in BATCH_SEEDS]
    batch = Batch(context)
    batch.set_graph_steps(16)
    sims = []
    for k, case in enumerate(cases):
        print(case.describe(), flush=True)
        sim = Simulation(context, case.scene)
        if k % 2:
            sim.set_accumulate_intensity(True)
        batch.add(sim, case.steps)
        sims.append(sim)
    batch.advance_by(max(case.steps for case in cases))
    batch.wait()
    for k, (case, sim) in enumerate(zip(cases, sims)):
        reference = oracle.Simulation(case.scene)
        if k % 2:
            reference.set_accumulate_intensity(True)
        reference.advance_by(case.steps)
        assert_matches_oracle(case, reference, sim, bool(k % 2))
    batch.close()
    for sim in sims:
        sim.close()


def test_many_random_scenes_in_random_chunks(context):
    # A wider sweep in one test: each draw runs one of the fused forms (one
    # step per launch, two on one wave, two on stacked waves) at a tiling
    # from its seed, advanced in chunks of random lengths (crossing injection
    # windows and odd leftover steps), with the intensity sum switched on
    # halfway for some. DIAPHANE_SWEEP_CASES widens it (3300 draws have run
    # clean on MI355X).
    import os

    count = int(os.environ.get("DIAPHANE_SWEEP_CASES", "300"))
    failures = []
    for seed in range(1000, 1000 + count):
        case = random_case(seed)
        rng = np.random.default_rng(90_000 + seed)
        form = int(rng.integers(3))
        if form == 0:
            rows, flags = int(rng.choice([4, 6, 8])), int(rng.choice([0, TILE_NONTEMPORAL, TILE_BUFFER]))
        elif form == 1:
            rows = int(rng.choice([6, 8, 10]))
            flags = TILE_TWO_STEP | int(rng.choice([0, TILE_BUFFER, TILE_NONTEMPORAL]))
        else:
            rows = int(rng.choice([4, 6]))
            flags = (TILE_TWO_STEP | TILE_STACK | (int(rng.choice([2, 4])) << 16) | int(rng.choice([0, TILE_BUFFER]))
                     | int(rng.choice([0, TILE_XCD_REMAP])))
        depth = int(rng.choice([0, 1, 3, 8, 64]))
        intensity = bool(rng.integers(2))
        chunks, left = [], case.steps
        while left:
            chunks.append(int(min(left, rng.choice([1, 2, 3, 5, 7, 64, 100]))))
            left -= chunks[-1]
        reference = oracle.Simulation(case.scene)
        device = Simulation(context, case.scene)
        device.set_tile(rows, depth, flags)
        for k, chunk in enumerate(chunks):
            if intensity and k == len(chunks) // 2:
                reference.set_accumulate_intensity(True)
                device.set_accumulate_intensity(True)
            reference.advance_by(chunk)
            device.advance_by(chunk)
        same = (bits_equal(reference.electric_all, device.read_electric())
                and bits_equal(reference.magnetic_all, device.read_magnetic())
                and (not intensity or np.array_equal(reference.intensity_f64(), device.read_intensity_f64())))
        device.close()
        if not same:
            failures.append(f"{case.describe()} rows {rows} depth {depth} flags {flags:#x} chunks {chunks}")
    print(f"{count} random scenes in random chunks, {len(failures)} differing", flush=True)
    assert not failures, "\n".join(failures[:10])


BATCH_SHAPES = [
    (8, 0, TILE_TWO_STEP | TILE_BUFFER | TILE_XCD_REMAP | (4 << 8)),  # the batch default
    (6, 16, TILE_TWO_STEP),
    (10, 3, TILE_TWO_STEP | TILE_NONTEMPORAL),
    (4, 8, TILE_TWO_STEP | TILE_STACK | (4 << 16) | TILE_BUFFER),
    (6, 0, TILE_TWO_STEP | TILE_STACK | (2 << 16)),
    (4, 0, TILE_NONTEMPORAL),  # one step per launch, persistent grid
    (6, 5, 16 | TILE_BUFFER),  # one step per launch, one wave per tile
    (8, 64, 0),
]


@pytest.mark.parametrize("shape", BATCH_SHAPES, ids=[f"{r}-{d}-{f:#x}" for r, d, f in BATCH_SHAPES])
def test_random_scenes_batched_on_every_batch_tiling(context, shape):
    # 24 draws in one batch at each tiling, to their own step counts, advanced
    # in chunks of random lengths, a third of them summing intensity.
    b = BATCH_SHAPES.index(shape)
    cases = [random_case(seed) for seed in range(5000 + 1000 * b, 5024 + 1000 * b)]
    batch = Batch(context)
    batch.set_tile(*shape)
    sims = []
    for k, case in enumerate(cases):
        sim = Simulation(context, case.scene)
        if k % 3 == 0:
            sim.set_accumulate_intensity(True)
        batch.add(sim, case.steps)
        sims.append(sim)
    rng = np.random.default_rng(b)
    steps, done = max(case.steps for case in cases), 0
    while done < steps:
        chunk = int(min(steps - done, rng.choice([1, 3, 17, 40, 100])))
        batch.advance_by(chunk)
        done += chunk
    batch.wait()
    for k, (case, sim) in enumerate(zip(cases, sims)):
        reference = oracle.Simulation(case.scene)
        if k % 3 == 0:
            reference.set_accumulate_intensity(True)
        reference.advance_by(case.steps)
        assert_matches_oracle(case, reference, sim, k % 3 == 0)
    batch.close()
    for sim in sims:
        sim.close()
