// Host-side scene resolution for the engine: grid spacing and grading,
// material coefficients, absorber profiles, waveforms, injections and
// rasterization. Everything here runs once at construction, except the
// per-step injection evaluation, which is O(sources) per step.
//
// Arithmetic follows the reference's f32 expressions operation for operation
// (compiled with -ffp-contract=off) so the resolved arrays are bit-identical
// to what kvark/diaphane's host code produces on the same libm.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/diaphane_amd.h"

namespace dpa {

// src/grid.rs:42-48
constexpr float kSpeedOfLight = 299792458.0f;
constexpr float kVacuumPermeability = 1.2566371e-6f;
constexpr float kVacuumPermittivity = 8.854188e-12f;
constexpr float kCourantLimit = 0.57735026f;
constexpr float kPi = 3.14159265358979323846f;

struct Error {
    int code;
    std::string message;
};

// One axis of cell widths (src/grid.rs:150-286).
struct Axis1D {
    std::vector<float> primary;  // cell widths, the H update
    std::vector<float> dual;     // centre-to-centre, the E update
    std::vector<float> corners;  // count + 1, f64-summed

    static Axis1D uniform(uint32_t count, float width);
    static Axis1D from_widths(std::vector<float> widths);
    uint32_t count() const { return static_cast<uint32_t>(primary.size()); }
    float length() const { return corners.back(); }
    float finest() const;
    float coarsest() const;
    float worst_ratio() const;
    float position(float coordinate) const;
    float coordinate(float position) const;
};

struct Grid {
    uint32_t extent[3] = {0, 0, 0};
    float courant = 0.5f;
    Axis1D axis[3];

    // Grid::build + Grid::validate (src/grid.rs:466-500, 747-773).
    static Grid resolve(const dp_grid_spec &spec);
    uint64_t cells() const {
        return static_cast<uint64_t>(extent[0]) * extent[1] * extent[2];
    }
    float reference_cell_size() const;
    float time_step() const;
    std::vector<float> gains(const std::vector<float> &widths) const;
    void cell_containing(const float position[3], uint32_t out[3]) const;
    // Grid::cell_lookup (src/grid.rs:635-646): per axis x, y, z, `samples`
    // fractional cell coordinates of positions spread evenly from one wall
    // to the other (samples >= 2).
    std::vector<float> cell_lookup(uint32_t samples) const;
    // The sample count gpu::Simulation::new picks (src/gpu.rs:225-227): four
    // per cell of the longest axis, within 256..8192.
    uint32_t lookup_samples() const;
    bool contains(const float position[3]) const;
    float coarsest() const;
};

struct Coefficients {
    float electric_gain, electric_loss, magnetic_gain, magnetic_loss;
};
Coefficients material_coefficients(const dp_material &m, float time_step);

// The six 1D loss profiles (src/boundary.rs:100-145).
struct Absorber {
    std::vector<float> integer[3];
    std::vector<float> half[3];
    float peak = 0.0f;
    static Absorber build(const Grid &grid, const dp_scene &scene);
};

float waveform_value(const dp_waveform &w, float time);
float waveform_dominant_frequency(const dp_waveform &w);

// Source::injections for one source (src/source.rs:280-403).
void source_injections(const dp_source &s, const Grid &grid, float time,
                       dp_injection out[2]);
// Injection::weight at a cell centre (src/source.rs:441-455).
float injection_weight(const dp_injection &inj, const Grid &grid,
                       const uint32_t coord[3]);

// Scene::material_indices (src/scene.rs:181-206), any integer width.
void rasterize(const dp_scene &scene, const Grid &grid, uint32_t *out);
void rasterize_u8(const dp_scene &scene, const Grid &grid, uint8_t *out);

// Scene::validate (src/scene.rs:254-369); empty string when valid.
std::string validate(const dp_scene &scene);

// A scene read from RON text, owning the arrays its dp_scene points into.
struct SceneFile {
    dp_scene scene{};
    std::vector<dp_refinement> refinements;
    std::vector<dp_material> materials;
    std::vector<dp_shape> shapes;
    std::vector<dp_source> sources;
    void bind();  // points `scene` at the vectors
};

// Scene::from_ron / to_ron (src/scene.rs:467-498): the RON text format.
std::unique_ptr<SceneFile> parse_scene_ron(const std::string &text);
std::string scene_to_ron(const dp_scene &scene);
bool scene_is_reversible(const dp_scene &scene);

// Thread-local last error, shared by every C entry point.
void set_error(const std::string &message);
const char *last_error();

}  // namespace dpa
