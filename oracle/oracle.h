/*
 * CPU oracle for the Yee step: a plain-C restatement of kvark/diaphane's
 * src/cpu.rs and of the scene layer it builds on (src/grid.rs,
 * src/material.rs, src/boundary.rs, src/source.rs, src/scene.rs).
 *
 * TEST INFRASTRUCTURE ONLY. Only the tests, the smoke check and the CPU
 * baseline of the benchmark load this library; the engine never links it.
 *
 * The input format is the engine's public scene description (dp_scene in
 * include/diaphane_amd.h); nothing else of the engine is used.
 */
#ifndef DIAPHANE_ORACLE_H
#define DIAPHANE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/diaphane_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

/* Last error message of this thread. */
const char *or_last_error(void);

/* Grid resolution (src/grid.rs:466-500): extent, primary widths x|y|z,
 * time step. */
int or_grid_resolve(const dp_grid_spec *grid, uint32_t extent_out[3],
                    float *widths_out, size_t capacity, float *time_step_out);
/* Dual widths x|y|z (src/grid.rs:201-208). */
int or_grid_dual(const dp_grid_spec *grid, float *out, size_t capacity);
/* Packed geometry (src/grid.rs:666-678): 3 sections × (nx+ny+nz). */
int or_grid_packed_geometry(const dp_grid_spec *grid, float *out,
                            size_t capacity);
/* Spacing::coordinate / position on one axis (src/grid.rs:253-286). */
float or_grid_coordinate(const dp_grid_spec *grid, int axis, float position);
float or_grid_position(const dp_grid_spec *grid, int axis, float coordinate);
/* Grid::cell_containing (src/grid.rs:710-714). */
int or_grid_cell_lookup(const dp_grid_spec *spec, uint32_t samples, float *out, size_t capacity);
int or_grid_cell_containing(const dp_grid_spec *grid, const float position[3],
                            uint32_t out[3]);
/* numerical_phase_velocity (src/grid.rs:966-1012). */
float or_numerical_phase_velocity(float courant, float cells_per_wavelength,
                                  const float direction[3]);

/* Material::coefficients (src/material.rs:100-122) for a whole table:
 * out[4*i + {0,1,2,3}] = electric_gain, electric_loss, magnetic_gain,
 * magnetic_loss. */
int or_material_coefficients(const dp_material *materials, uint32_t count,
                             float time_step, float *out);

/* AbsorbingProfile::packed (src/boundary.rs:100-212):
 * [int x | int y | int z | half x | half y | half z]. */
int or_absorber_packed(const dp_scene *scene, float *out, size_t capacity);

/* Waveform::evaluate (src/source.rs:74-115). */
float or_waveform_evaluate(const dp_waveform *waveform, float time);

/* Source::injections for every source, two per source. */
int or_scene_injections(const dp_scene *scene, float time, dp_injection *out,
                        size_t capacity);
/* Injection::weight at a cell centre (src/source.rs:441-455). */
float or_injection_weight(const dp_scene *scene, const dp_injection *inj,
                          const uint32_t coord[3]);

/* Scene::material_indices (src/scene.rs:181-206). */
int or_scene_material_indices(const dp_scene *scene, uint32_t *out,
                              size_t count);

/* cpu::Simulation (src/cpu.rs:64-760). threads <= 0 uses all cores; the
 * sweeps split over z planes, which leaves every cell's arithmetic and
 * therefore every result bit-identical to the single-threaded run. */
or_sim *or_sim_create(const dp_scene *scene);
void or_sim_destroy(or_sim *sim);
void or_sim_set_threads(or_sim *sim, int threads);
uint64_t or_sim_cells(const or_sim *sim);
void or_sim_extent(const or_sim *sim, uint32_t out[3]);
uint64_t or_sim_step_count(const or_sim *sim);
float or_sim_time(const or_sim *sim);
void or_sim_advance(or_sim *sim, uint64_t steps);
int or_sim_is_reversible(const or_sim *sim);
int or_sim_reverse(or_sim *sim, uint64_t steps);
void or_sim_reset(or_sim *sim);
void or_sim_set_accumulate_intensity(or_sim *sim, int on);
uint64_t or_sim_accumulated_steps(const or_sim *sim);
/* Direct views (electric_mut/magnetic_mut, src/cpu.rs:145-151): 3·N floats
 * each, component-major. */
float *or_sim_electric(or_sim *sim);
float *or_sim_magnetic(or_sim *sim);
/* The reference's f32 running sum, and the same per-step f32 sums
 * accumulated in f64 (the engine's accumulator). */
float *or_sim_intensity(or_sim *sim);
double *or_sim_intensity_f64(or_sim *sim);
/* Steppable::restore (src/cpu.rs:736-752). */
int or_sim_restore(or_sim *sim, const float *electric, const float *magnetic,
                   size_t count, uint64_t step);
/* energy (src/cpu.rs:654-687). */
void or_sim_energy(const or_sim *sim, double *electric, double *magnetic);
int or_sim_is_finite(const or_sim *sim);

/* The f64 track: the forward update with double fields and arithmetic on
 * the same (promoted) f32 tables and injections. begin copies the current
 * f32 state; advance moves only the f64 fields. For stating the f32
 * rounding error of the update per workload. */
int or_sim_f64_begin(or_sim *sim);
int or_sim_f64_advance(or_sim *sim, uint64_t steps);
double *or_sim_electric_f64(or_sim *sim);
double *or_sim_magnetic_f64(or_sim *sim);
uint64_t or_sim_step_count_f64(const or_sim *sim);

#ifdef __cplusplus
}
#endif
#endif
