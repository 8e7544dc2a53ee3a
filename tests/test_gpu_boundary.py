"""The rest of gpu::Simulation's surface on the device: the renderer's
lookup buffer (src/gpu.rs:225-228, 345-351), reversal's refusals
(src/gpu.rs:499-509), Timeline::clear on a ring moved to another grid
(src/timeline.rs:157-160), and injection lists longer than the kernels'
32-entry mask (a scene of 40 point sources, and one of 34 plane waves
whose H and E rows both exceed it)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from diaphane_amd import Axis, Boundary, Extent, Grid, Material, Refinement, Scene, Shape, Source, Waveform, _lib
from diaphane_amd.batch import Batch
from diaphane_amd.gpu import BUFFER_LOOKUP, TILE_BUFFER, TILE_TWO_STEP, TILE_XCD_REMAP, Simulation
from diaphane_amd.timeline import DeviceTimeline, Seek

from test_gpu_device_buffers import device_read

pytestmark = pytest.mark.gpu


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------------------
# lookup_buffer / lookup_samples
# ---------------------------------------------------------------------------
def lookup_scenes():
    graded = Grid.graded([0.064, 0.04, 0.048], 1000.0, [Refinement.across(Axis.X, 0.0, 0.0064, 0.25e-3)])
    return [Scene.photon(Extent(40, 32, 30)), Scene.cavity(Extent(90, 70, 3000 // 40)), Scene.on_grid(graded)]


@pytest.mark.parametrize("index", range(3))
def test_the_lookup_buffer_is_the_oracle_table(context, index):
    scene = lookup_scenes()[index]
    sim = Simulation(context, scene)
    samples = sim.lookup_samples()
    longest = max(scene.grid.extent.as_array())
    assert samples == min(max(4 * longest, 256), 8192)
    table = device_read(sim.device_ptr(BUFFER_LOOKUP), 3 * samples, np.float32)
    assert bits_equal(table, oracle.cell_lookup(scene.grid, samples))
    # the buffer is the grid's, not the state's: stepping leaves it alone
    sim.advance_by(5)
    sim.wait()
    assert bits_equal(device_read(sim.device_ptr(BUFFER_LOOKUP), 3 * samples, np.float32), table)
    sim.close()


# ---------------------------------------------------------------------------
# reverse_by's refusals: a status and the reference's message, nothing changed
# ---------------------------------------------------------------------------
LOSSY = "this scene is lossy"


def pulse_scene(extent: Extent) -> Scene:
    grid = Grid.new(extent, 1e-3)
    return Scene(grid).with_source(Source.point([0.0, 0.0, 0.0], Axis.Z,
                                                Waveform.ricker(grid.frequency_for_resolution(20.0))))


def conductor_scene() -> Scene:
    scene = pulse_scene(Extent.cube(32)).with_boundary(Boundary.PEC)
    metal = scene.materials.push(Material.PERFECT_CONDUCTOR)
    scene.shapes.append(Shape.Block([8e-3, 0.0, 0.0], [4e-3] * 3, metal))
    return scene


def assert_refused(sim, steps, code, message):
    step = sim.step_count()
    e, h = sim.read_electric(), sim.read_magnetic()
    with pytest.raises(_lib.DiaphaneError, match=message) as caught:
        sim.reverse_by(steps)
    assert caught.value.code == code
    assert sim.step_count() == step
    assert bits_equal(sim.read_electric(), e) and bits_equal(sim.read_magnetic(), h)


@pytest.mark.parametrize("make", [lambda: pulse_scene(Extent.cube(32)), conductor_scene],
                         ids=["absorber", "conductor"])
def test_reversal_refuses_a_lossy_scene(context, make):
    # src/cpu.rs:972-998 and src/gpu.rs:503-508
    sim = Simulation(context, make())
    assert not sim.is_reversible()
    sim.advance_by(3)
    sim.set_accumulate_intensity(True)
    sim.advance_by(2)
    assert_refused(sim, 1, _lib.DP_ERR_LOSSY, LOSSY)
    assert sim.accumulated_steps() == 2  # a refused reversal clears nothing
    sim.reverse_by(0)  # no steps: nothing to refuse (src/gpu.rs:500-502)
    reference = oracle.Simulation(make())
    reference.advance_by(5)
    assert bits_equal(sim.read_electric(), reference.electric_all)
    sim.close()


def test_reversal_refuses_to_pass_the_start_and_then_unwinds(context):
    # src/gpu.rs:509; then src/cpu.rs:953-970 (reverse to an empty domain)
    scene = pulse_scene(Extent.cube(24)).with_boundary(Boundary.PEC)
    sim = Simulation(context, scene)
    assert sim.is_reversible()
    sim.advance_by(120)
    excited = sim.energy().total()
    assert_refused(sim, 121, _lib.DP_ERR_PAST_START, "cannot step back past the start")
    sim.reverse_by(120)
    assert sim.step_count() == 0
    assert sim.energy().total() < excited * 1e-8
    reference = oracle.Simulation(scene)
    reference.advance_by(120)
    reference.reverse_by(120)
    assert bits_equal(sim.read_electric(), reference.electric_all)
    assert bits_equal(sim.read_magnetic(), reference.magnetic_all)
    sim.close()


# ---------------------------------------------------------------------------
# Timeline::clear when the scene underneath changes
# ---------------------------------------------------------------------------
def test_a_cleared_ring_records_and_seeks_on_another_grid(context):
    first, second = Scene.photon(Extent.cube(24)), Scene.photon(Extent(40, 28, 26))
    timeline = DeviceTimeline(10, 4)
    a = Simulation(context, first)
    for _ in range(5):
        a.advance_by(7)
        timeline.observe(a)
    assert timeline.keyframe_count() >= 3
    b = Simulation(context, second)
    b.advance_by(40)  # past the ring's newest frame (35), so observe records
    with pytest.raises(_lib.DiaphaneError) as caught:  # the ring is still bound to the first grid
        timeline.observe(b)
    assert caught.value.code == _lib.DP_ERR_SIZE
    timeline.clear()
    assert timeline.keyframe_count() == 0 and timeline.bytes() == 0
    for _ in range(6):
        b.advance_by(7)
        timeline.observe(b)
    assert timeline.bytes() == timeline.keyframe_count() * 24 * second.grid.extent.total()
    outcome = timeline.seek(b, 70)
    assert outcome == Seek("from_keyframe", 9, 61) and b.step_count() == 70
    reference = oracle.Simulation(second)
    reference.advance_by(70)
    assert bits_equal(b.read_electric(), reference.electric_all)
    assert bits_equal(b.read_magnetic(), reference.magnetic_all)
    a.close()
    b.close()
    timeline.close()


# ---------------------------------------------------------------------------
# more than 32 injections a phase
# ---------------------------------------------------------------------------
def many_points() -> Scene:
    # 140 × 50 × 40 cells: one-step tiles are 64 wide, two-step tiles own 60
    # columns from x = −2, so the points straddle both kinds of seam, and
    # y rows 1..6 put them on row seams of 2-, 4- and 6-row ownership.
    scene = Scene.photon(Extent(140, 50, 40))
    grid = scene.grid
    frequency = grid.frequency_for_resolution(20.0)
    scene.sources = []
    xs = [0, 1, 57, 58, 59, 62, 63, 64, 65, 117, 118, 119, 127, 128, 139]
    k = 0
    for i in range(40):
        x = xs[i % len(xs)]
        y = 1 + (i * 7) % 48
        z = (i * 5) % 40
        at = [oracle.position(grid, 0, x + 0.5), oracle.position(grid, 1, y + 0.5), oracle.position(grid, 2, z + 0.5)]
        pol = Axis(i % 3)
        wave = Waveform.ricker(frequency * (0.8 + 0.01 * i))
        scene.sources.append(Source.point(at, pol, wave).with_amplitude(0.5 + 0.03 * i))
        k += 1
    return scene


def many_plane_waves() -> Scene:
    # 34 plane waves over a PEC box: each contributes an H row and an E row,
    # so both phases carry 34 injections per step.
    grid = Grid.new(Extent(70, 44, 36), 1e-3)
    scene = Scene(grid, Boundary.PEC)
    frequency = grid.frequency_for_resolution(24.0)
    size = grid.size()
    for i in range(34):
        travel = Axis(i % 3)
        pol = travel.next() if i % 2 else travel.prev()
        offset = (-0.4 + 0.8 * ((i * 11) % 34) / 34.0) * size[int(travel)]
        wave = Waveform.gaussian_pulse(frequency * (0.9 + 0.005 * i), 2.0)
        scene.sources.append(Source.plane_wave(travel, offset, pol, wave).with_amplitude(0.05))
    return scene


def test_the_scenes_exceed_the_mask():
    for scene, magnetic, electric in ((many_points(), 0, 40), (many_plane_waves(), 34, 34)):
        reference = oracle.Simulation(scene)
        time = 30 * scene.grid.time_step()
        injections = [q for q in oracle.scene_injections(scene, time) if q.value != 0.0]
        assert sum(1 for q in injections if q.magnetic) == magnetic
        assert sum(1 for q in injections if not q.magnetic) == electric
        del reference


PAIR = (10, 0, TILE_TWO_STEP | TILE_BUFFER | TILE_XCD_REMAP)


@pytest.mark.parametrize("make", [many_points, many_plane_waves], ids=["points", "plane_waves"])
@pytest.mark.parametrize("tile", [None, PAIR, (6, 3, TILE_TWO_STEP)], ids=["one_step", "two_step", "two_step_6"])
def test_long_injection_lists_match_the_oracle(context, make, tile):
    scene = make()
    steps = 47
    sim = Simulation(context, scene)
    if tile:
        sim.set_tile(*tile)
    sim.advance_by(steps)
    reference = oracle.Simulation(scene)
    reference.advance_by(steps)
    assert np.abs(reference.electric_all).max() > 0.0
    assert bits_equal(sim.read_electric(), reference.electric_all)
    assert bits_equal(sim.read_magnetic(), reference.magnetic_all)
    sim.close()


def test_long_injection_lists_in_a_batch(context):
    scenes = [many_points(), many_plane_waves(), Scene.photon(Extent(33, 29, 27))]
    batch = Batch(context)
    batch.set_graph_steps(8)
    members = [Simulation(context, s) for s in scenes]
    for sim in members:
        batch.add(sim)
    batch.advance_by(21)
    batch.wait()
    for scene, sim in zip(scenes, members):
        reference = oracle.Simulation(scene)
        reference.advance_by(21)
        assert bits_equal(sim.read_electric(), reference.electric_all)
        assert bits_equal(sim.read_magnetic(), reference.magnetic_all)
    batch.close()
