/*
 * diaphane_amd — MI355X (gfx950) engine for diaphane's 3D Yee FDTD solver.
 *
 * C ABI. Plain structs, pointers and sizes; no C++ or torch types cross it.
 * Every entry point that can fail returns an int status (DP_OK == 0) and
 * leaves a thread-local message in dp_last_error(). The messages keep the
 * wording of the reference's panics so a binding can raise the same text.
 *
 * Citations are to kvark/diaphane (paths relative to the crate root).
 *
 * Ownership: the caller owns every host array it passes in (copied at the
 * call) or out (filled at the call). The library owns device memory. A dp_sim
 * is externally synchronized; separate sims may be driven from separate
 * threads.
 */
#ifndef DIAPHANE_AMD_H
#define DIAPHANE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                             */
/* ------------------------------------------------------------------------ */
enum {
    DP_OK = 0,
    DP_ERR_INVALID = -1,     /* bad argument / invalid grid or scene        */
    DP_ERR_LOSSY = -2,       /* reversal of a lossy scene (cpu.rs:253-260)  */
    DP_ERR_PAST_START = -3,  /* reversing past step 0 (gpu.rs:509)          */
    DP_ERR_SIZE = -4,        /* buffer length mismatch (gpu.rs:665-669)     */
    DP_ERR_TIMEOUT = -5,     /* wait timed out (gpu.rs:616-622)             */
    DP_ERR_DEVICE = -6,      /* HIP runtime failure                         */
    DP_ERR_UNSUPPORTED = -7, /* outside what this build supports            */
    DP_ERR_BUDGET = -8       /* device-memory budget exhausted (serving)    */
};

/* ------------------------------------------------------------------------ */
/* Scene description (the input format). Mirrors the reference's Scene,    */
/* Grid/GridSpec, Material, Shape, Source and Waveform field for field.     */
/* Positions are metres from the domain centre, as in the reference.        */
/* ------------------------------------------------------------------------ */
enum { DP_AXIS_X = 0, DP_AXIS_Y = 1, DP_AXIS_Z = 2 };

/* Refinement, src/grid.rs:313-336. size[a] <= 0 opts axis a out. */
typedef struct dp_refinement {
    float center[3];
    float size[3];
    float cell_size;
} dp_refinement;

/* GridSpec, src/grid.rs:381-392: the recipe, resolved by grade(). */
typedef struct dp_grid_spec {
    uint32_t extent[3]; /* cells the domain has at cell_size            */
    float cell_size;    /* coarse cell size, metres                     */
    float courant;      /* S = c·Δt/Δ_ref, (0, 0.57735026]              */
    float max_ratio;    /* growth cap between neighbours (default 1.15) */
    const dp_refinement *refinements;
    uint32_t refinement_count;
} dp_grid_spec;

/* Material, src/material.rs:26-39. conductivity = +inf is a PEC. */
typedef struct dp_material {
    float relative_permittivity;
    float relative_permeability;
    float conductivity;
    float magnetic_loss;
} dp_material;

/* Shape, src/scene.rs:32-56. Later shapes overwrite earlier ones. */
enum { DP_SHAPE_BLOCK = 0, DP_SHAPE_SPHERE = 1, DP_SHAPE_SLAB = 2 };
typedef struct dp_shape {
    uint32_t kind;
    uint32_t material;
    float center[3];  /* block, sphere  */
    float size[3];    /* block          */
    float radius;     /* sphere         */
    uint32_t axis;    /* slab           */
    float offset;     /* slab           */
    float thickness;  /* slab           */
} dp_shape;

/* Waveform, src/source.rs:31-52, plus a differentiated Gaussian.
 *   RICKER:          frequency = peak_frequency, delay
 *   GAUSSIAN_PULSE:  frequency = center_frequency, cycles = duration_cycles
 *                    (FWHM in carrier periods), delay
 *   CONTINUOUS_WAVE: frequency, cycles = ramp_cycles
 *   DIFF_GAUSSIAN:   width = σ in seconds, delay; value
 *                    −u·exp(½(1 − u²)) with u = (t − delay)/σ (peak |1| at
 *                    u = ∓1); dominant frequency 1/(2πσ). No counterpart in
 *                    the reference; both the oracle and the engine
 *                    evaluate it on the host like every other waveform. */
enum {
    DP_WAVE_RICKER = 0,
    DP_WAVE_GAUSSIAN_PULSE = 1,
    DP_WAVE_CONTINUOUS_WAVE = 2,
    DP_WAVE_DIFF_GAUSSIAN = 3
};
typedef struct dp_waveform {
    uint32_t kind;
    float frequency;
    float cycles;
    float delay;
    float width;
} dp_waveform;

/* Source + SourceShape, src/source.rs:136-197. */
enum { DP_SOURCE_POINT = 0, DP_SOURCE_SHEET = 1, DP_SOURCE_PLANE_WAVE = 2 };
typedef struct dp_source {
    uint32_t shape;
    uint32_t polarization; /* driven E component                     */
    float at[3];           /* point                                  */
    float velocity[3];     /* point, m/s                             */
    uint32_t axis;         /* sheet normal / plane-wave travel axis  */
    float offset;          /* sheet, plane wave                      */
    float waist;           /* sheet                                  */
    dp_waveform waveform;
    float amplitude;
} dp_source;

/* Boundary, src/boundary.rs:53-75. */
enum { DP_BOUNDARY_PEC = 0, DP_BOUNDARY_ABSORBING = 1 };

/* Scene, src/scene.rs:116-122. materials[0] must be vacuum. */
typedef struct dp_scene {
    dp_grid_spec grid;
    uint32_t boundary;
    uint32_t absorber_thickness;   /* cells, absorbing only       */
    float target_reflection;       /* absorbing only, in (0, 1)   */
    const dp_material *materials;
    uint32_t material_count;
    const dp_shape *shapes;
    uint32_t shape_count;
    const dp_source *sources;
    uint32_t source_count;
} dp_scene;

/* ------------------------------------------------------------------------ */
/* Host-side scene resolution (no device needed).                           */
/* ------------------------------------------------------------------------ */

/* Scene files in the reference's RON format: Scene::from_ron and to_ron
 * (src/scene.rs:467-498; the GridSpec checks of src/grid.rs:894-935 come
 * back as errors). dp_scene_from_ron parses `text` (NUL-terminated) into a
 * scene the returned file owns; a syntax error names its line and column.
 * The scene is not validated: call dp_scene_validate. */
typedef struct dp_scene_file dp_scene_file;
int dp_scene_from_ron(const char *text, dp_scene_file **out);
/* The parsed scene, valid until dp_scene_file_destroy. */
const dp_scene *dp_scene_file_scene(const dp_scene_file *file);
void dp_scene_file_destroy(dp_scene_file *file);
/* Writes `scene` as RON that dp_scene_from_ron reads back to the same
 * values. *length gets the text's length (without the NUL); with out ==
 * NULL nothing is written, otherwise capacity must hold text and NUL
 * (DP_ERR_SIZE if not). */
int dp_scene_to_ron(const dp_scene *scene, char *out, size_t capacity, size_t *length);

/* Message of the last failed call on this thread ("" if none). */
const char *dp_last_error(void);

/* Library version string. */
const char *dp_version(void);

/* Page-locked host memory for repeated readbacks (dp_sim_read_*,
 * dp_sim_write_fields): copies to and from it run at full PCIe rate with no
 * staging. No counterpart in the reference, whose readbacks map its own
 * staging buffers. dp_host_free(NULL) does nothing. */
int dp_host_alloc(size_t bytes, void **out);
void dp_host_free(void *ptr);

/* Resolves a GridSpec: Grid::build (src/grid.rs:466-500, grade() at
 * :791-888). extent_out gets the resolved cell counts; widths_out (may be
 * NULL) gets the primary widths of x, then y, then z
 * (extent[0]+extent[1]+extent[2] floats, widths_capacity checked);
 * time_step_out (may be NULL) gets Grid::time_step (src/grid.rs:597-599). */
int dp_grid_resolve(const dp_grid_spec *grid, uint32_t extent_out[3],
                    float *widths_out, size_t widths_capacity,
                    float *time_step_out);

/* Grid::cell_lookup (src/grid.rs:635-646): the renderer's world-to-cell
 * table. Per axis x, then y, then z, `samples` fractional cell coordinates
 * (Spacing::coordinate, src/grid.rs:269-279) of the positions
 * (s/(samples−1) − ½)·length, s = 0..samples−1, all in f32. `out` holds
 * 3·samples floats; samples >= 2. */
int dp_grid_cell_lookup(const dp_grid_spec *grid, uint32_t samples, float *out, size_t capacity);
/* The lookup sample count gpu::Simulation::new picks (src/gpu.rs:225-227):
 * 4 × the longest axis's cell count, clamped to 256..8192. */
int dp_grid_lookup_samples(const dp_grid_spec *grid, uint32_t *samples);

/* Scene::validate (src/scene.rs:254-369) for the recoverable problems;
 * DP_OK or DP_ERR_INVALID with the reason in dp_last_error(). */
int dp_scene_validate(const dp_scene *scene);

/* Scene::material_indices (src/scene.rs:181-206), widened to u32 so it can
 * be compared with the reference's Vec<u32>. count = resolved cell count. */
int dp_scene_material_indices(const dp_scene *scene, uint32_t *out,
                              size_t count);

/* One injection as Source::injections resolves it (src/source.rs:409-439),
 * already sign-scaled. weight_kind 0: every cell weight 1; 1: Gaussian
 * apodization (center, inverse_waist_squared) evaluated on the host. */
typedef struct dp_injection {
    uint32_t origin[3];
    uint32_t extent[3];
    uint32_t component;
    uint32_t magnetic;
    float value;
    float center[3];
    float inverse_waist_squared;
} dp_injection;

/* The tables the kernels read, as the engine resolves them on the host:
 * per material (electric_gain, electric_loss, magnetic_gain, magnetic_loss)
 * (Material::coefficients, src/material.rs:100-122; 4·material_count
 * floats), and per axis 4·(nx+ny+nz) floats laid out
 * [primary gains x|y|z][dual gains x|y|z][integer absorber x|y|z]
 * [half absorber x|y|z] (Grid::packed_geometry sections 0-1,
 * src/grid.rs:666-678; AbsorbingProfile::packed, src/boundary.rs:203-212).
 * Either output may be NULL. */
int dp_scene_tables(const dp_scene *scene, float *coefficients,
                    size_t coefficient_capacity, float *axis_table,
                    size_t axis_capacity);

/* Source::injections(grid, time) for every source of the scene, in order,
 * two per source (src/source.rs:280-403). `out` holds 2·source_count. */
int dp_scene_injections(const dp_scene *scene, float time, dp_injection *out,
                        size_t capacity);

/* ------------------------------------------------------------------------ */
/* Device engine: the counterpart of gpu::Simulation (src/gpu.rs:159-799). */
/* ------------------------------------------------------------------------ */
typedef struct dp_engine dp_engine;
typedef struct dp_sim dp_sim;

/* headless_context (src/gpu.rs:805-817): opens HIP device `device`. */
int dp_engine_create(int device, dp_engine **out);
void dp_engine_destroy(dp_engine *engine);
/* Name of the device's ISA, e.g. "gfx950". */
int dp_engine_arch(dp_engine *engine, char *out, size_t capacity);

/* gpu::Simulation::new (src/gpu.rs:190-340). Validates the grid, resolves
 * the scene on the host (materials, coefficients, absorber, geometry,
 * apodization weights) and uploads it; fields start at zero. */
int dp_sim_create(dp_engine *engine, const dp_scene *scene, dp_sim **out);
/* Drop (src/gpu.rs:779-799): waits, then frees every device buffer. */
void dp_sim_destroy(dp_sim *sim);

/* grid().extent, step_count, time (src/gpu.rs:353-364), field_bytes and
 * bytes_per_step (src/gpu.rs:377-388: 12·N and 48·N). */
int dp_sim_extent(const dp_sim *sim, uint32_t out[3]);
uint64_t dp_sim_step_count(const dp_sim *sim);
float dp_sim_time(const dp_sim *sim);
float dp_sim_time_step(const dp_sim *sim);
uint64_t dp_sim_field_bytes(const dp_sim *sim);
uint64_t dp_sim_bytes_per_step(const dp_sim *sim);
int dp_sim_is_reversible(const dp_sim *sim);

/* advance_by (src/gpu.rs:393-484): enqueues `steps` full steps on the
 * sim's stream and returns without waiting. The call's injection table is
 * copied through one of four event-guarded slots, so back-to-back calls queue
 * on the stream; a call blocks only while the device is still four
 * table-carrying calls behind. */
int dp_sim_advance(dp_sim *sim, uint64_t steps);
/* reverse_by (src/gpu.rs:499-576): DP_ERR_LOSSY / DP_ERR_PAST_START. */
int dp_sim_reverse(dp_sim *sim, uint64_t steps);
/* wait (src/gpu.rs:616-622); timeout_ms 0 means the reference's 120 s. */
int dp_sim_wait(dp_sim *sim, uint32_t timeout_ms);
/* reset (src/gpu.rs:625-637): zero E, H, intensity; step = 0. */
int dp_sim_reset(dp_sim *sim);

/* set_accumulate_intensity / accumulated_steps (src/gpu.rs:584-594). The
 * device accumulator is f64 per cell; each step adds the f32 sum
 * ((Ex²+Ey²)+Ez²) exactly as cpu.rs:186-196 forms it. */
int dp_sim_set_accumulate_intensity(dp_sim *sim, int on);
uint64_t dp_sim_accumulated_steps(const dp_sim *sim);

/* read_electric / read_magnetic (src/gpu.rs:644-656): 3·N floats,
 * component-major, x fastest. Synchronous. */
int dp_sim_read_electric(dp_sim *sim, float *out, size_t count);
int dp_sim_read_magnetic(dp_sim *sim, float *out, size_t count);
/* An overlapped read_electric / read_magnetic (no counterpart in the
 * reference, whose readbacks block: src/gpu.rs:644-656). begin_read returns
 * at once: E (which = DP_BUFFER_ELECTRIC) or H (DP_BUFFER_MAGNETIC) as of
 * this call, after every step queued so far, is copied on the device to a
 * staging buffer of the sim and from there to `out` (3·N floats) on a
 * stream of its own, while later advances run. Two readbacks can be in
 * flight per sim, each with its own staging buffer (12 bytes a cell of
 * device memory each, allocated at first use and kept); their host copies
 * land in the order they were begun, and begin_read with two in flight
 * first waits for the older. `out` must stay valid and untouched until its
 * readback has landed; page-locked memory (dp_host_alloc) lets the host copy
 * overlap the steps. wait_reads waits until at most `in_flight` readbacks
 * are in flight (in_flight 1 after a begin_read: every earlier one has
 * landed, so a caller can use the previous buffer of two while the newest
 * copy runs); finish_read waits for all of them (wait_reads with 0). Both
 * poll without sleeping for their first 2 ms (timeout 0: 120 s;
 * DP_ERR_TIMEOUT). Destroying the sim waits for every readback in flight. */
int dp_sim_begin_read(dp_sim *sim, int which, float *out, size_t count);
int dp_sim_wait_reads(dp_sim *sim, uint32_t in_flight, uint32_t timeout_ms);
int dp_sim_finish_read(dp_sim *sim, uint32_t timeout_ms);
/* read_intensity (src/gpu.rs:702-705): N floats (the f64 sums rounded). */
int dp_sim_read_intensity(dp_sim *sim, float *out, size_t count);
int dp_sim_read_intensity_f64(dp_sim *sim, double *out, size_t count);

/* Steppable::restore (src/gpu.rs:756-762) and seeding: uploads E and H
 * (3·N floats each), sets the step, clears the intensity sum. */
int dp_sim_write_fields(dp_sim *sim, const float *electric,
                        const float *magnetic, size_t count, uint64_t step);

/* The viewer's auto-range reduction (measure_peak, examples/common/
 * render.rs:382-419 with shaders/volume.wgsl:122-132, 376-378, 473-493), so
 * a renderer that binds dp_sim_device_ptr buffers gets its exposure without
 * a readback of the fields. peak[0]: max over cells of the larger of ½|E|²
 * and ½|H|² (each sum ((0+x²)+y²)+z² in f32); peak[1..3]: max |E_x|, |E_y|,
 * |E_z|; peak[4..6]: max |H_x|, |H_y|, |H_z|; peak[7]: max of the intensity
 * sum (rounded to f32) over the accumulated step count, 0 with no sum.
 * Synchronous. */
int dp_sim_measure_peak(dp_sim *sim, float peak[8]);
/* The same eight maxima, bit for bit, from a call that allocates nothing.
 * No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). The per-block partials and the final maxima live in a fixed
 * file-scope device array of 128 KiB + 32 B per GPU, held until the process
 * ends; its device address is taken once per device by dp_engine_create.
 * The reduction finishes on the device, so no host buffer is allocated
 * either. DP_ERR_SIZE if a grid would need more than its 4096 partial
 * slots (the peak kernel never uses more). Calls that overlap on one GPU
 * share the array and race: serialize them. Synchronous. */
int dp_sim_measure_peak_noalloc(dp_sim *sim, float peak[8]);

/* cpu::Simulation::energy (src/cpu.rs:654-687) computed on the device:
 * per-component volumes from the primary/dual widths, f64 accumulation. */
int dp_sim_energy(dp_sim *sim, double *electric, double *magnetic);

/* electric_buffer / magnetic_buffer / intensity_buffer (src/gpu.rs:
 * 367-374, 597-599): device pointers, borrowed. The engine ping-pongs
 * E and H between two allocations each, so the pointer names the buffer
 * holding the current state and stays valid until the next advance,
 * reverse, reset or write. DP_BUFFER_INTENSITY is the f64 sum. For a
 * renderer that binds f32 like the reference's intensity_buffer,
 * DP_BUFFER_INTENSITY_F32 rounds the sum to f32 into a buffer of its own,
 * ordered on the sim's stream at this call: the values dp_sim_read_intensity
 * returns, as of this call. */
/* DP_BUFFER_LOOKUP is lookup_buffer (src/gpu.rs:345-347): the grid's
 * world-to-cell table, 3·dp_sim_lookup_samples floats as dp_grid_cell_lookup
 * gives them, uploaded once at dp_sim_create and valid until dp_sim_destroy. */
enum { DP_BUFFER_ELECTRIC = 0, DP_BUFFER_MAGNETIC = 1, DP_BUFFER_INTENSITY = 2,
       DP_BUFFER_MATERIAL_INDEX = 3, DP_BUFFER_INTENSITY_F32 = 4, DP_BUFFER_LOOKUP = 5 };
int dp_sim_device_ptr(dp_sim *sim, int which, void **out);
/* lookup_samples (src/gpu.rs:349-351): samples per axis in DP_BUFFER_LOOKUP. */
uint32_t dp_sim_lookup_samples(const dp_sim *sim);

/* Kernel path for the forward step. FUSED (default): one z-marching pass
 * per step doing H, H-injections, E, E-injections and intensity.
 * TWO_PASS: the reference's dispatch shape (H kernel, injections, E kernel,
 * injections, intensity). Both give bit-identical fields. */
enum { DP_KERNEL_FUSED = 0, DP_KERNEL_TWO_PASS = 1, DP_KERNEL_FUSED_OVERLAP = 2 };
/* DP_KERNEL_FUSED uses 64-aligned tiles whose x halo is fetched by the edge
 * lanes; DP_KERNEL_FUSED_OVERLAP uses 62-wide tiles whose halo columns are
 * ordinary lanes. */
enum { DP_TILE_XCD_REMAP = 1, DP_TILE_NONTEMPORAL = 2, DP_TILE_BUFFER = 4, DP_TILE_TWO_STEP = 8,
       DP_TILE_STACK = 32 };
int dp_sim_set_kernel(dp_sim *sim, int kernel);
/* Fused-kernel tiling: rows per wave along y (4, 6, 8, 10 or 12), z-slab
 * depth in planes, and flags:
 *   DP_TILE_XCD_REMAP     renumbers workgroups so each XCD takes runs of G
 *                         consecutive blocks of tiles (G in bits 8..15 of
 *                         flags; 0 means 8);
 *   DP_TILE_NONTEMPORAL   streams the field stores past the caches;
 *   DP_TILE_BUFFER        addresses the fields through buffer resources with
 *                         32-bit lane offsets (used when the grid's offsets
 *                         fit);
 *   DP_TILE_TWO_STEP      advances two steps per launch: the first step's
 *                         fields stay in registers and the second consumes
 *                         them one z-plane behind, so each pair of steps
 *                         reads and writes the fields once. Waves own 60
 *                         columns and rows − 4 of their rows (6, 8 or 10;
 *                         other values select 8); an odd step left over in
 *                         an advance runs single.
 *   DP_TILE_STACK         with DP_TILE_TWO_STEP: the waves of a workgroup
 *                         (W in bits 16..19: 2 or 4, 0 meaning 4) stack
 *                         along y on one tile of W·rows register rows
 *                         (rows 4 or 6; other values select 6), of which
 *                         W·rows − 4 are owned; neighbouring waves trade
 *                         their edge rows through LDS at three barriers per
 *                         plane instead of recomputing them.
 * A new sim of at least 768 Ki cells whose offsets fit buffer addressing
 * starts with two steps per launch on stacked tiles (four waves of 4 rows,
 * buffer addressing, XCD runs of 16); smaller ones with one step (4 rows,
 * non-temporal stores).
 * 0 for rows selects 4; 0 for depth selects a size-aware default (the power
 * of two at or below the number of x-y tiles / 50, within 1..32, or / 25
 * within 2..32 for two steps per launch: short slabs give a small grid
 * enough waves, long ones amortize a large grid's halo planes). flags < 0
 * keeps the current flags. With DP_TILE_TWO_STEP, the last step of an odd
 * count runs one step per launch on 4 rows with the size-aware depth and
 * this tiling's store, addressing and XCD flags. Results never depend on any
 * of it. */
int dp_sim_set_tile(dp_sim *sim, int rows, int depth, int flags);
/* Autotuner for DP_KERNEL_FUSED: times `trials` launches of every rows
 * (4, 6, 8 one step per launch with non-temporal stores; 6, 8, 10 two steps
 * per launch with either store kind and XCD runs of 8; 4 or 6 rows in 2 or 4
 * stacked waves with cached stores and XCD runs of 16) × depth (1 to 64 in powers of two, up
 * to twice the grid's z extent) × addressing (pointer, buffer where it fits)
 * candidate on this sim's grid, times the four fastest again with 3·trials
 * launches in two alternating rounds, then three quarters and one and a half
 * of the winner's depth the same way, keeps the fastest per step and reports
 * it (any out pointer may be null). A sim keeps one tiling for plain steps and one
 * for intensity-accumulating steps; the autotuner sets the one for the
 * sim's current mode (dp_sim_set_tile sets both). The state is unchanged:
 * trial launches read the current fields, apply the injections of the sim's
 * next two steps, write only the idle half of each ping-pong pair, and
 * accumulate into a scratch buffer. Synchronous. */
int dp_sim_autotune(dp_sim *sim, uint32_t trials, int *rows, int *depth, int *flags, float *ms_per_step);

/* HIP-event timing on the sim's stream: record event `slot` (0..7) now;
 * elapsed milliseconds between two recorded slots (waits for the later). */
int dp_sim_record_event(dp_sim *sim, int slot);
int dp_sim_elapsed_ms(dp_sim *sim, int from_slot, int to_slot, float *ms);

/* Bandwidth (GB/s) of the step's access mix — E, H and the u8 material
 * read, E and H written, 49 bytes per cell — streamed with no stencil on
 * this sim's own buffers, averaged over `iterations` launches. The state is
 * left unchanged (only the idle half of each ping-pong pair is written).
 * The practical ceiling for the step's byte count. No counterpart in the
 * reference: a measurement aid. */
int dp_sim_stream_probe(dp_sim *sim, int iterations, double *gbs);

/* Bandwidth (GB/s, 49 bytes per cell) of the fused kernel's own access
 * pattern with no stencil: the same 64-aligned tiles of `rows` rows (4, 6 or
 * 8; 0 takes the sim's plain tiling) marching z-slabs of `depth` planes (0:
 * the sim's), the same row and x-halo loads and owned-row stores, on this
 * sim's buffers. The state is left unchanged. The ceiling for that tiling's
 * memory traffic. No counterpart in the reference: a measurement aid. */
int dp_sim_tile_probe(dp_sim *sim, int rows, int depth, int iterations, double *gbs);

/* Keyframes in device memory: Timeline (src/timeline.rs:76-212) whose
 * snapshots (E and H, 24 bytes per cell) stay in HBM. A ring of `capacity`
 * keyframes, one per `interval` steps (Timeline::new, :113-125); the grid
 * size is fixed by the first keyframe (DP_ERR_SIZE for another grid later).
 * Recording and restoring are device-to-device copies on the sim's stream.
 * Not for batch members (they are advanced by their batch). */
typedef struct dp_timeline dp_timeline;
enum { DP_SEEK_UNCHANGED = 0, DP_SEEK_FORWARD = 1, DP_SEEK_FROM_KEYFRAME = 2, DP_SEEK_FROM_START = 3 };
int dp_timeline_create(uint64_t interval, uint32_t capacity, dp_timeline **out);
void dp_timeline_destroy(dp_timeline *timeline);
/* Timeline::observe (:127-148): records the sim's state if it has crossed
 * into a new interval past the newest keyframe. */
int dp_timeline_observe(dp_timeline *timeline, dp_sim *sim);
/* Timeline::seek (:167-211): forward by stepping; backward by restoring the
 * newest keyframe at or before `target` (which clears the intensity sum, as
 * a restore does) and replaying, or from a reset when none is. kind gets
 * DP_SEEK_*, replayed the steps replayed, at the keyframe's step (each out
 * pointer may be null). */
int dp_timeline_seek(dp_timeline *timeline, dp_sim *sim, uint64_t target, int *kind, uint64_t *replayed,
                     uint64_t *at);
/* Timeline::clear (src/timeline.rs:157-160), for when the scene underneath
 * has changed: drops every keyframe, frees their device memory and forgets
 * the grid size, so the next keyframe may come from a sim of another size.
 * Waits for pending copies into or out of the ring. */
int dp_timeline_clear(dp_timeline *timeline);
/* keyframe_count, earliest (0 when there is none) and bytes (:150-165). */
int dp_timeline_info(const dp_timeline *timeline, uint32_t *keyframes, uint64_t *earliest, uint64_t *bytes);

/* Forward-step launches enqueued so far (fused or H-kernel launches). */
uint64_t dp_sim_launch_count(const dp_sim *sim);

/* ------------------------------------------------------------------------ */
/* Batches: many independent sims advanced by one launch per step           */
/* ------------------------------------------------------------------------ */
/* The reference runs one gpu::Simulation per scene, N steps per command
 * buffer (src/gpu.rs:393-484); it has no batched form. A dp_batch advances
 * its member sims together: every launch does one step of every active
 * member over a packed table of tiles, so small grids share the device
 * instead of each under-filling it. Each member's fields are bit-identical
 * to the same sim advanced alone.
 *
 * A member keeps every dp_sim read/write/reset/energy entry point; its work
 * is ordered on the batch's stream. dp_sim_advance and dp_sim_reverse refuse
 * a member. Destroying a member sim removes it from its batch; destroying a
 * batch releases its members (their state stays as it is). A batch and its
 * members are externally synchronized together. A member's field component
 * must fit in 4 GiB (cells ≤ 2^30). */
typedef struct dp_batch dp_batch;
int dp_batch_create(dp_engine *engine, dp_batch **out);
void dp_batch_destroy(dp_batch *batch);
int dp_batch_add(dp_batch *batch, dp_sim *sim);
int dp_batch_remove(dp_batch *batch, dp_sim *sim);
/* Step count at which a member stops; members without one run with every
 * advance. A member at its target sits out further launches. */
int dp_batch_set_target(dp_batch *batch, dp_sim *sim, uint64_t step);
/* Enqueues `steps` launches (fewer if every member reaches its target),
 * each advancing every member below its target by one step. Asynchronous. */
int dp_batch_advance(dp_batch *batch, uint64_t steps);
int dp_batch_wait(dp_batch *batch, uint32_t timeout_ms);
/* Launches per captured HIP graph (default 32; 0 launches each kernel
 * directly). Results never depend on it. */
int dp_batch_set_graph_steps(dp_batch *batch, uint32_t steps);
/* Rows per wave (4, 6 or 8; with DP_TILE_TWO_STEP 6, 8 or 10; 0 = default
 * 4, or 8 with DP_TILE_TWO_STEP, 4 with DP_TILE_STACK), z-slab depth (0, the default: sized to the membership, the
 * single-sim rule over all members' x-y tiles together, at most 16, or 32
 * for two steps) and
 * flags (< 0 keeps the current ones): DP_TILE_NONTEMPORAL, DP_TILE_BUFFER
 * (used when every member's offsets fit), DP_BATCH_WAVE_PER_TILE, which
 * launches one wave per tile instead of the persistent grid whose waves
 * stride over the tile list, and DP_TILE_TWO_STEP: each launch advances
 * every member two steps (one wave per tile; a member's last step of an odd
 * count runs alone in its launch). With DP_TILE_TWO_STEP, DP_TILE_XCD_REMAP
 * (run length G in bits 8..15, 0 meaning 8) gives each XCD runs of G
 * consecutive workgroups of the tile list; the one-step batch kernels ignore
 * it. DP_TILE_STACK (with DP_TILE_TWO_STEP; waves in bits 16..19, 2 or 4)
 * stacks the waves of a workgroup along y on one tile, as for a single sim
 * (rows 4 or 6). The default is two steps per launch, 8 rows, DP_TILE_BUFFER,
 * cached stores, XCD runs of 4. The autotuner keeps the XCD bits as they
 * are. */
enum { DP_BATCH_WAVE_PER_TILE = 16 };
int dp_batch_set_tile(dp_batch *batch, int rows, int depth, int flags);
/* Autotuner: times `trials` launches of every grid form × addressing ×
 * rows × depth candidate (persistent, one wave per tile, or two steps per
 * launch with either store kind, or two steps on 2 or 4 stacked waves with
 * buffer addressing; pointer or buffer; rows 4, 6, 8, or 6, 8, 10 for two
 * steps, 4 or 6 stacked; depth 8, 16, 32, 64) on the members as they are,
 * times the four fastest again with 3·trials launches in two alternating
 * rounds, keeps the fastest per step and reports rows, depth and time (any
 * out pointer may be null; dp_batch_tiling reports the form). Member state
 * is unchanged: trial launches write only the idle half of each ping-pong
 * pair. Synchronous. */
int dp_batch_autotune(dp_batch *batch, uint32_t trials, int *rows, int *depth, float *ms_per_step);
/* The tiling in use (any out pointer may be null). */
int dp_batch_tiling(const dp_batch *batch, int *rows, int *depth, int *flags);
uint32_t dp_batch_size(const dp_batch *batch);
uint64_t dp_batch_cells(const dp_batch *batch);
uint64_t dp_batch_launch_count(const dp_batch *batch);
int dp_batch_record_event(dp_batch *batch, int slot);
int dp_batch_elapsed_ms(dp_batch *batch, int from_slot, int to_slot, float *ms);

/* ------------------------------------------------------------------------ */
/* Serving: a request loop with continuous batching                         */
/* ------------------------------------------------------------------------ */
/* No counterpart in the reference (one gpu::Simulation per caller,
 * src/gpu.rs:190). A dp_server owns a batch, a loop thread and host worker
 * threads. Clients submit solves from any thread; a request is admitted
 * when its device footprint (49 bytes per cell plus a small fixed part)
 * fits the HBM budget, built by a worker while the device keeps stepping
 * the others, and joins the batch between chunks of launches. A finished
 * request's sim stays resident, readable through the dp_sim entry points,
 * until dp_server_release. Results are bit-identical to the same scene run
 * alone through dp_sim_advance. */
typedef struct dp_server dp_server;
typedef struct dp_server_config {
    uint64_t hbm_budget_bytes; /* 0: 200 GiB                              */
    uint32_t max_resident;     /* admitted, unfinished requests (0: 1024) */
    uint32_t chunk_steps;      /* launches between joins/retires (0: 32)  */
    uint32_t host_threads;     /* sim builders (0: 4)                     */
    int32_t one_at_a_time;     /* 1: run requests alone, in arrival order,
                                  through dp_sim_advance (the baseline)   */
    int32_t rows, depth;       /* batch tiling (0: defaults)              */
    uint64_t chunk_cells;      /* batched mode: members step in order of
                                  least remaining work (cells × steps
                                  left) until this many cells step in a
                                  chunk; the others wait that chunk, and
                                  none waits more than 64 chunks in a row.
                                  0: 128M cells; UINT64_MAX: every member
                                  steps every chunk                       */
} dp_server_config;
enum { DP_REQUEST_QUEUED = 0, DP_REQUEST_PREPARING = 1, DP_REQUEST_READY = 2, DP_REQUEST_RUNNING = 3,
       DP_REQUEST_DONE = 4 };
typedef struct dp_request_info {
    int32_t state;  /* DP_REQUEST_*                                   */
    int32_t status; /* DP_OK, or the failure once state is DONE       */
    uint64_t steps, step, cells, footprint_bytes;
    /* seconds since the server was created; 0 until reached */
    double submitted_s, admitted_s, started_s, completed_s;
} dp_request_info;
/* busy_s: loop time spent advancing and waiting for chunks; kernel_s: the
 * device time of those chunks' launches (first launch to last, by events);
 * cell_updates: cells × steps advanced, over all requests. */
typedef struct dp_server_stats {
    uint32_t queued, preparing, ready, running;
    uint64_t used_bytes, budget_bytes, chunks;
    double busy_s, uptime_s;
    double kernel_s;
    uint64_t cell_updates;
} dp_server_stats;
int dp_server_create(dp_engine *engine, const dp_server_config *config, dp_server **out);
/* Stops the loop; every request still held is dropped. */
void dp_server_destroy(dp_server *server);
/* Copies the scene; DP_ERR_BUDGET if the request alone exceeds the budget. */
int dp_server_submit(dp_server *server, const dp_scene *scene, uint64_t steps, uint64_t *ticket);
int dp_server_query(dp_server *server, uint64_t ticket, dp_request_info *info);
/* Blocks until the request is done; its failure status if it failed. A
 * ticket must not be released while another thread waits on it. */
int dp_server_wait(dp_server *server, uint64_t ticket, uint32_t timeout_ms);
/* The finished request's sim, borrowed until dp_server_release. */
int dp_server_result(dp_server *server, uint64_t ticket, dp_sim **sim);
/* Frees a finished request and its device memory (returns its budget). */
int dp_server_release(dp_server *server, uint64_t ticket);
int dp_server_statistics(dp_server *server, dp_server_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* DIAPHANE_AMD_H */
