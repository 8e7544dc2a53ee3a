// Host-side scene resolution. See scene.hpp.
#include "scene.hpp"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <thread>

namespace dpa {

namespace {
thread_local std::string g_last_error;

[[noreturn]] void raise(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void raise(int code, const char *fmt, ...) {
    char buffer[512];
    va_list args;
    va_start(args, fmt);
    std::vsnprintf(buffer, sizeof buffer, fmt, args);
    va_end(args);
    throw Error{code, buffer};
}

// Runs body(z) for z in [0, n) on the host's cores, split into contiguous
// chunks. Results never depend on the split: each z is independent.
template <typename F>
void parallel_z(uint32_t n, F body) {
    unsigned workers = std::max(1u, std::min(std::thread::hardware_concurrency(), 64u));
    workers = std::min<unsigned>(workers, std::max<uint32_t>(1, n));
    if (workers <= 1) {
        for (uint32_t z = 0; z < n; ++z) body(z);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(workers);
    for (unsigned w = 0; w < workers; ++w) {
        uint32_t begin = static_cast<uint32_t>(uint64_t(n) * w / workers);
        uint32_t end = static_cast<uint32_t>(uint64_t(n) * (w + 1) / workers);
        pool.emplace_back([=] {
            for (uint32_t z = begin; z < end; ++z) body(z);
        });
    }
    for (auto &t : pool) t.join();
}
}  // namespace

void set_error(const std::string &message) { g_last_error = message; }
const char *last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Axis1D (src/grid.rs:150-286)
// ---------------------------------------------------------------------------
static void fill_dual(Axis1D &s) {
    // dual[i] = ½(w[i−1] + w[i]), first cell mirrored (src/grid.rs:201-208).
    s.dual.resize(s.primary.size());
    for (size_t i = 0; i < s.primary.size(); ++i) {
        size_t prev = i == 0 ? 0 : i - 1;
        s.dual[i] = 0.5f * (s.primary[prev] + s.primary[i]);
    }
}

Axis1D Axis1D::uniform(uint32_t count, float width) {
    if (count < 2) raise(DP_ERR_INVALID, "an axis needs at least 2 cells, got %u", count);
    Axis1D s;
    s.primary.assign(count, width);
    s.corners.resize(count + 1);
    // Multiplied in f64, not accumulated (src/grid.rs:164-175).
    for (uint32_t i = 0; i <= count; ++i)
        s.corners[i] = static_cast<float>(static_cast<double>(i) * static_cast<double>(width));
    fill_dual(s);
    return s;
}

Axis1D Axis1D::from_widths(std::vector<float> widths) {
    if (widths.size() < 2)
        raise(DP_ERR_INVALID, "an axis needs at least 2 cells, got %zu", widths.size());
    for (float w : widths)
        if (!(w > 0.0f && std::isfinite(w)))
            raise(DP_ERR_INVALID, "cell widths must be positive and finite");
    Axis1D s;
    s.primary = std::move(widths);
    s.corners.resize(s.primary.size() + 1);
    double running = 0.0;  // rounding once per corner (src/grid.rs:178-188)
    s.corners[0] = 0.0f;
    for (size_t i = 0; i < s.primary.size(); ++i) {
        running += static_cast<double>(s.primary[i]);
        s.corners[i + 1] = static_cast<float>(running);
    }
    fill_dual(s);
    return s;
}

float Axis1D::finest() const {
    float v = INFINITY;
    for (float w : primary) v = std::fmin(v, w);
    return v;
}

float Axis1D::coarsest() const {
    float v = -INFINITY;
    for (float w : primary) v = std::fmax(v, w);
    return v;
}

float Axis1D::worst_ratio() const {
    float worst = 1.0f;
    for (size_t i = 0; i + 1 < primary.size(); ++i)
        worst = std::fmax(worst, std::fmax(primary[i + 1] / primary[i], primary[i] / primary[i + 1]));
    return worst;
}

float Axis1D::position(float coordinate) const {
    const uint32_t last = count() - 1;
    float floor0 = std::fmax(std::floor(coordinate), 0.0f);
    uint32_t index = floor0 >= static_cast<float>(last) ? last : static_cast<uint32_t>(floor0);
    float offset = corners[index] + (coordinate - static_cast<float>(index)) * primary[index];
    return offset - 0.5f * length();
}

float Axis1D::coordinate(float position) const {
    float offset = position + 0.5f * length();
    // partition_point(corner <= offset) − 1, clamped to the last cell.
    auto past = std::upper_bound(corners.begin(), corners.end(), offset);
    size_t index = static_cast<size_t>(past - corners.begin());
    index = index == 0 ? 0 : index - 1;
    index = std::min<size_t>(index, primary.size() - 1);
    return static_cast<float>(index) + (offset - corners[index]) / primary[index];
}

// grade(), src/grid.rs:791-888.
static Axis1D grade(int axis, float length, float base, const dp_refinement *all,
                    uint32_t all_count, float max_ratio) {
    std::vector<const dp_refinement *> refs;
    for (uint32_t i = 0; i < all_count; ++i)
        if (all[i].size[axis] > 0.0f) refs.push_back(&all[i]);
    if (refs.empty())
        return Axis1D::uniform(static_cast<uint32_t>(std::fmax(std::round(length / base), 2.0f)), base);

    float finest = base;
    for (auto *r : refs) finest = std::fmin(finest, r->cell_size);
    if (!(finest > 0.0f && std::isfinite(finest)))
        raise(DP_ERR_INVALID, "refinement cell sizes must be positive and finite, got %g", finest);
    const double wanted = std::ceil(static_cast<double>(length / finest * 4.0f));
    const size_t samples = static_cast<size_t>(std::clamp(wanted, 64.0, double(1u << 20)));
    const float step = length / static_cast<float>(samples);

    std::vector<float> target(samples + 1, base);
    for (size_t i = 0; i < target.size(); ++i) {
        const float position = static_cast<float>(i) * step - 0.5f * length;
        for (auto *r : refs)
            if (std::fabs(position - r->center[axis]) <= 0.5f * r->size[axis])
                target[i] = std::fmin(target[i], r->cell_size);
    }
    // Two-sided Lipschitz bound: a distance transform, exact in two passes.
    const float slope = (max_ratio - 1.0f) * step;
    for (size_t i = 1; i < target.size(); ++i) target[i] = std::fmin(target[i], target[i - 1] + slope);
    for (size_t i = target.size() - 1; i-- > 0;) target[i] = std::fmin(target[i], target[i + 1] + slope);

    std::vector<float> widths;
    float offset = 0.0f;
    while (offset < length) {
        size_t index = std::min(static_cast<size_t>(offset / step), target.size() - 1);
        widths.push_back(target[index]);
        offset += target[index];
        if (widths.size() > (1u << 16))
            raise(DP_ERR_INVALID,
                  "grading the %d axis ran to %zu cells; the refinements are finer than the "
                  "domain can carry", axis, widths.size());
    }
    // Re-apply the cap to the cells themselves, both directions.
    for (size_t i = 1; i < widths.size(); ++i) widths[i] = std::fmin(widths[i], widths[i - 1] * max_ratio);
    for (size_t i = widths.size() - 1; i-- > 0;) widths[i] = std::fmin(widths[i], widths[i + 1] * max_ratio);

    float total = -0.0f;
    for (float w : widths) total += w;
    if (widths.size() > 2) {
        const float without = total - widths.back();
        if (std::fabs(length - without) < std::fabs(total - length)) {
            widths.pop_back();
            total = without;
        }
    }
    const float scale = length / total;
    for (float &w : widths) w *= scale;
    return Axis1D::from_widths(std::move(widths));
}

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------
Grid Grid::resolve(const dp_grid_spec &spec) {
    if (!(spec.cell_size > 0.0f && std::isfinite(spec.cell_size)))
        raise(DP_ERR_INVALID, "cell size must be positive and finite, got %g", spec.cell_size);
    if (!(spec.max_ratio > 1.0f && std::isfinite(spec.max_ratio)))
        raise(DP_ERR_INVALID, "the growth cap must exceed 1, got %g", spec.max_ratio);
    if (spec.refinement_count > 0 && spec.refinements == nullptr)
        raise(DP_ERR_INVALID, "refinements pointer is null");
    Grid g;
    for (int a = 0; a < 3; ++a) {
        if (spec.refinement_count == 0)
            g.axis[a] = Axis1D::uniform(spec.extent[a], spec.cell_size);
        else
            g.axis[a] = grade(a, static_cast<float>(spec.extent[a]) * spec.cell_size, spec.cell_size,
                              spec.refinements, spec.refinement_count, spec.max_ratio);
        g.extent[a] = g.axis[a].count();
    }
    g.courant = spec.courant;
    if (!(g.courant > 0.0f && g.courant <= kCourantLimit))
        raise(DP_ERR_INVALID, "Courant number %g is outside (0, %g]; the scheme would diverge",
              static_cast<double>(g.courant), static_cast<double>(kCourantLimit));
    return g;
}

float Grid::reference_cell_size() const {
    float inverse_squares = -0.0f;
    for (int a = 0; a < 3; ++a) {
        float f = axis[a].finest();
        inverse_squares += 1.0f / (f * f);
    }
    return std::sqrt(3.0f / inverse_squares);
}

float Grid::time_step() const { return courant * reference_cell_size() / kSpeedOfLight; }

std::vector<float> Grid::gains(const std::vector<float> &widths) const {
    const float travel = kSpeedOfLight * time_step();
    std::vector<float> out(widths.size());
    for (size_t i = 0; i < widths.size(); ++i) out[i] = travel / widths[i];
    return out;
}

void Grid::cell_containing(const float position[3], uint32_t out[3]) const {
    for (int a = 0; a < 3; ++a) {
        float c = std::fmax(std::floor(axis[a].coordinate(position[a])), 0.0f);
        uint32_t last = extent[a] - 1;
        out[a] = c >= static_cast<float>(last) ? last : static_cast<uint32_t>(c);
    }
}

std::vector<float> Grid::cell_lookup(uint32_t samples) const {
    std::vector<float> out;
    out.reserve(size_t(3) * samples);
    const float last = static_cast<float>(samples - 1);
    for (int a = 0; a < 3; ++a) {
        const float length = axis[a].length();
        for (uint32_t s = 0; s < samples; ++s) {
            const float fraction = static_cast<float>(s) / last;
            out.push_back(axis[a].coordinate((fraction - 0.5f) * length));
        }
    }
    return out;
}

uint32_t Grid::lookup_samples() const {
    const uint32_t longest = std::max(extent[0], std::max(extent[1], extent[2]));
    const uint64_t four = 4ull * longest;
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(four, 256), 8192));
}

bool Grid::contains(const float position[3]) const {
    for (int a = 0; a < 3; ++a)
        if (!(std::fabs(position[a]) <= 0.5f * axis[a].length())) return false;
    return true;
}

float Grid::coarsest() const {
    float v = -INFINITY;
    for (int a = 0; a < 3; ++a) v = std::fmax(v, axis[a].coarsest());
    return v;
}

// ---------------------------------------------------------------------------
// Materials (src/material.rs:100-122)
// ---------------------------------------------------------------------------
Coefficients material_coefficients(const dp_material &m, float time_step) {
    if (std::isinf(m.conductivity))  // PEC: the loss = 1 limit
        return {0.0f, 1.0f, 1.0f / m.relative_permeability, 0.0f};
    const float half_step = 0.5f * time_step;
    return {1.0f / m.relative_permittivity,
            half_step * m.conductivity / (kVacuumPermittivity * m.relative_permittivity),
            1.0f / m.relative_permeability,
            half_step * m.magnetic_loss / (kVacuumPermeability * m.relative_permeability)};
}

// ---------------------------------------------------------------------------
// Absorber (src/boundary.rs:100-145)
// ---------------------------------------------------------------------------
Absorber Absorber::build(const Grid &grid, const dp_scene &scene) {
    Absorber p;
    uint32_t thickness = 0;
    float target = 1.0f;
    if (scene.boundary == DP_BOUNDARY_ABSORBING) {
        thickness = scene.absorber_thickness;
        target = scene.target_reflection;
    } else if (scene.boundary != DP_BOUNDARY_PEC) {
        raise(DP_ERR_INVALID, "unknown boundary kind %u", scene.boundary);
    }
    for (int a = 0; a < 3; ++a) {
        p.integer[a].assign(grid.extent[a], 0.0f);
        p.half[a].assign(grid.extent[a], 0.0f);
    }
    if (thickness == 0) return p;
    for (int a = 0; a < 3; ++a)
        if (!(2 * uint64_t(thickness) < grid.extent[a]))
            raise(DP_ERR_INVALID, "absorbing layers of %u cells do not fit in axis %d of %u cells",
                  thickness, a, grid.extent[a]);
    if (!(target > 0.0f && target < 1.0f))
        raise(DP_ERR_INVALID, "target reflection %g must lie in (0, 1)", static_cast<double>(target));
    const float order = 3.0f;
    const float t = static_cast<float>(thickness);
    const float peak = -(order + 1.0f) * std::log(target) * grid.courant / (4.0f * t);
    auto profile = [&](float position, uint32_t n) {
        const float outer = static_cast<float>(n - 1);
        const float low = (t - position) / t;
        const float high = (t - (outer - position)) / t;
        const float depth = std::clamp(std::fmax(low, high), 0.0f, 1.0f);
        return peak * std::pow(depth, order);
    };
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < grid.extent[a]; ++i) {
            p.integer[a][i] = profile(static_cast<float>(i), grid.extent[a]);
            p.half[a][i] = profile(static_cast<float>(i) + 0.5f, grid.extent[a]);
            p.peak = std::fmax(p.peak, std::fmax(p.integer[a][i], p.half[a][i]));
        }
    return p;
}

// ---------------------------------------------------------------------------
// Waveforms and injections (src/source.rs)
// ---------------------------------------------------------------------------
float waveform_value(const dp_waveform &w, float time) {
    switch (w.kind) {
    case DP_WAVE_RICKER: {
        const float arg = kPi * w.frequency * (time - w.delay);
        const float arg2 = arg * arg;
        return (1.0f - 2.0f * arg2) * std::exp(-arg2);
    }
    case DP_WAVE_GAUSSIAN_PULSE: {
        const float fwhm_to_sigma = 0.4246609f;  // 1/(2√(2 ln 2))
        const float sigma = fwhm_to_sigma * w.cycles / w.frequency;
        const float tau = time - w.delay;
        const float envelope = std::exp(-0.5f * (tau / sigma) * (tau / sigma));
        return envelope * std::sin(2.0f * kPi * w.frequency * tau);
    }
    case DP_WAVE_CONTINUOUS_WAVE: {
        if (time <= 0.0f) return 0.0f;
        const float ramp = w.cycles / w.frequency;
        const float envelope = time >= ramp ? 1.0f : 0.5f * (1.0f - std::cos(kPi * time / ramp));
        return envelope * std::sin(2.0f * kPi * w.frequency * time);
    }
    case DP_WAVE_DIFF_GAUSSIAN: {
        const float u = (time - w.delay) / w.width;
        return -u * std::exp(0.5f * (1.0f - u * u));
    }
    default:
        return 0.0f;
    }
}

float waveform_dominant_frequency(const dp_waveform &w) {
    if (w.kind == DP_WAVE_DIFF_GAUSSIAN) return 1.0f / (2.0f * kPi * w.width);
    return w.frequency;
}

void source_injections(const dp_source &s, const Grid &grid, float time, dp_injection out[2]) {
    const float value = s.amplitude * waveform_value(s.waveform, time);
    dp_injection base{};
    switch (s.shape) {
    case DP_SOURCE_POINT: {
        float at[3];
        for (int a = 0; a < 3; ++a) at[a] = s.at[a] + s.velocity[a] * time;
        grid.cell_containing(at, base.origin);
        base.extent[0] = base.extent[1] = base.extent[2] = 1;
        base.component = s.polarization;
        // Gone means silent (src/source.rs:296-300).
        base.value = grid.contains(at) ? value : 0.0f;
        out[0] = base;
        out[1] = base;
        out[1].value = 0.0f;
        return;
    }
    case DP_SOURCE_SHEET: {
        const uint32_t a = s.axis;
        float at[3] = {0.0f, 0.0f, 0.0f};
        at[a] = s.offset;
        uint32_t cell[3];
        grid.cell_containing(at, cell);
        for (int t = 0; t < 3; ++t) base.extent[t] = grid.extent[t];
        base.origin[a] = cell[a];
        base.extent[a] = 1;
        base.inverse_waist_squared = 1.0f / (s.waist * s.waist);
        base.component = s.polarization;
        base.value = value;
        out[0] = base;
        out[1] = base;
        out[1].value = 0.0f;
        return;
    }
    default: {  // TFSF plane wave (src/source.rs:334-401)
        const uint32_t travel = s.axis;
        float at[3] = {0.0f, 0.0f, 0.0f};
        at[travel] = s.offset;
        uint32_t cell[3];
        grid.cell_containing(at, cell);
        const uint32_t plane = std::max<uint32_t>(cell[travel], 1);
        const float spacing = grid.axis[travel].finest();
        const float dt = grid.time_step();
        const float gain = kSpeedOfLight * dt / spacing;
        const float omega = 2.0f * kPi * waveform_dominant_frequency(s.waveform);
        const float half_k_spacing = std::asin(std::sin(0.5f * omega * dt) / gain);
        const float retard = half_k_spacing / omega;
        const uint32_t pol = s.polarization;
        const uint32_t h_axis = 3 - travel - pol;
        const float h_sign = ((pol + 1) % 3 == travel) ? -1.0f : 1.0f;
        auto drive = [&](float t) { return s.amplitude * waveform_value(s.waveform, t); };
        for (int t = 0; t < 3; ++t) base.extent[t] = grid.extent[t];
        base.extent[travel] = 1;
        dp_injection e = base, h = base;
        e.origin[travel] = plane;
        e.component = pol;
        e.magnetic = 0;
        e.value = gain * drive(time - 0.5f * dt + retard);
        h.origin[travel] = plane - 1;
        h.component = h_axis;
        h.magnetic = 1;
        h.value = h_sign * gain * drive(time - dt);
        out[0] = e;
        out[1] = h;
        return;
    }
    }
}

float injection_weight(const dp_injection &inj, const Grid &grid, const uint32_t coord[3]) {
    if (inj.inverse_waist_squared == 0.0f) return 1.0f;
    float r2 = 0.0f;
    for (int a = 0; a < 3; ++a)
        if (inj.extent[a] > 1) {
            const float centre = grid.axis[a].position(static_cast<float>(coord[a]) + 0.5f);
            const float off = centre - inj.center[a];
            r2 += off * off;
        }
    return std::exp(-r2 * inj.inverse_waist_squared);
}

// ---------------------------------------------------------------------------
// Rasterization (src/scene.rs:60-81, 181-206)
// ---------------------------------------------------------------------------
namespace {
struct ShapeMask {
    const dp_shape *shape;
    // Per-axis membership for the separable shapes (block: all three
    // axes; slab: its own axis). A sphere is tested per cell.
    std::vector<uint8_t> inside[3];
};

template <typename T>
void rasterize_into(const dp_scene &scene, const Grid &grid, T *out) {
    const uint32_t nx = grid.extent[0], ny = grid.extent[1], nz = grid.extent[2];
    std::vector<float> centre[3];
    for (int a = 0; a < 3; ++a) {
        centre[a].resize(grid.extent[a]);
        for (uint32_t i = 0; i < grid.extent[a]; ++i)
            centre[a][i] = grid.axis[a].position(static_cast<float>(i) + 0.5f);
    }
    std::vector<ShapeMask> masks(scene.shape_count);
    for (uint32_t s = 0; s < scene.shape_count; ++s) {
        const dp_shape &sh = scene.shapes[s];
        masks[s].shape = &sh;
        for (int a = 0; a < 3; ++a) {
            auto &in = masks[s].inside[a];
            in.assign(grid.extent[a], 1);
            for (uint32_t i = 0; i < grid.extent[a]; ++i) {
                const float p = centre[a][i];
                if (sh.kind == DP_SHAPE_BLOCK)
                    in[i] = std::fabs(p - sh.center[a]) <= 0.5f * sh.size[a];
                else if (sh.kind == DP_SHAPE_SLAB && sh.axis == static_cast<uint32_t>(a))
                    in[i] = std::fabs(p - sh.offset) <= 0.5f * sh.thickness;
            }
        }
    }
    parallel_z(nz, [&](uint32_t z) {
        std::vector<const ShapeMask *> active;
        active.reserve(masks.size());
        for (uint32_t y = 0; y < ny; ++y) {
            T *row = out + (static_cast<uint64_t>(z) * ny + y) * nx;
            active.clear();
            for (const auto &m : masks) {
                const dp_shape &sh = *m.shape;
                if (sh.kind == DP_SHAPE_SPHERE) {
                    // (dx²+dy²)+dz² >= dy² and >= dz², so either exceeding r²
                    // already rules the row out, exactly.
                    const float dy = centre[1][y] - sh.center[1], dz = centre[2][z] - sh.center[2];
                    const float rr = sh.radius * sh.radius;
                    if (dy * dy > rr || dz * dz > rr) continue;
                } else if (!(m.inside[1][y] && m.inside[2][z])) {
                    continue;
                }
                active.push_back(&m);
            }
            if (active.empty()) {
                std::fill(row, row + nx, T(0));
                continue;
            }
            for (uint32_t x = 0; x < nx; ++x) {
                T index = 0;  // vacuum
                for (const ShapeMask *m : active) {
                    const dp_shape &sh = *m->shape;
                    bool hit;
                    if (sh.kind == DP_SHAPE_SPHERE) {
                        const float dx = centre[0][x] - sh.center[0];
                        const float dy = centre[1][y] - sh.center[1];
                        const float dz = centre[2][z] - sh.center[2];
                        const float d2 = (dx * dx + dy * dy) + dz * dz;
                        hit = d2 <= sh.radius * sh.radius;
                    } else {
                        hit = m->inside[0][x];
                    }
                    if (hit) index = static_cast<T>(sh.material);  // later shapes win
                }
                row[x] = index;
            }
        }
    });
}
}  // namespace

void rasterize(const dp_scene &scene, const Grid &grid, uint32_t *out) {
    rasterize_into(scene, grid, out);
}

void rasterize_u8(const dp_scene &scene, const Grid &grid, uint8_t *out) {
    rasterize_into(scene, grid, out);
}

// ---------------------------------------------------------------------------
// Scene::validate (src/scene.rs:254-369)
// ---------------------------------------------------------------------------
static bool is_vacuum(const dp_material &m) {
    return m.relative_permittivity == 1.0f && m.relative_permeability == 1.0f &&
           m.conductivity == 0.0f && m.magnetic_loss == 0.0f;
}

bool scene_is_reversible(const dp_scene &scene) {
    if (scene.boundary != DP_BOUNDARY_PEC) return false;
    for (uint32_t i = 0; i < scene.material_count; ++i)
        if (!(scene.materials[i].conductivity == 0.0f && scene.materials[i].magnetic_loss == 0.0f))
            return false;
    return true;
}

std::string validate(const dp_scene &scene) {
    char buffer[512];
    Grid grid;
    try {
        grid = Grid::resolve(scene.grid);
    } catch (const Error &e) {
        return e.message;
    }
    if (scene.material_count == 0 || scene.materials == nullptr)
        return "the material table is empty; slot 0 must be vacuum";
    if (!is_vacuum(scene.materials[0]))
        return "material 0 must be vacuum: every unpainted cell wears index 0";
    if (scene.boundary == DP_BOUNDARY_ABSORBING) {
        for (int a = 0; a < 3; ++a)
            if (2 * uint64_t(scene.absorber_thickness) >= grid.extent[a]) {
                std::snprintf(buffer, sizeof buffer,
                              "absorbing layers %u cells thick meet in the middle of axis %d, which "
                              "is only %u cells across; thin the layer or raise the resolution",
                              scene.absorber_thickness, a, grid.extent[a]);
                return buffer;
            }
        if (!(scene.target_reflection > 0.0f && scene.target_reflection < 1.0f)) {
            std::snprintf(buffer, sizeof buffer, "target reflection %g must lie in (0, 1)",
                          static_cast<double>(scene.target_reflection));
            return buffer;
        }
    }
    for (uint32_t s = 0; s < scene.shape_count; ++s)
        if (scene.shapes[s].material >= scene.material_count) {
            std::snprintf(buffer, sizeof buffer, "shape references material %u but the table has %u",
                          scene.shapes[s].material, scene.material_count);
            return buffer;
        }
    // Resolution against the cells each material actually occupies.
    if (scene.source_count > 0) {
        std::vector<uint32_t> indices(grid.cells());
        rasterize(scene, grid, indices.data());
        std::vector<float> coarsest(scene.material_count, 0.0f);
        const uint32_t nx = grid.extent[0], ny = grid.extent[1], nz = grid.extent[2];
        for (uint32_t z = 0; z < nz; ++z)
            for (uint32_t y = 0; y < ny; ++y) {
                const float slab = std::fmax(grid.axis[1].primary[y], grid.axis[2].primary[z]);
                const uint64_t row = (uint64_t(z) * ny + y) * nx;
                for (uint32_t x = 0; x < nx; ++x) {
                    const float cell = std::fmax(slab, grid.axis[0].primary[x]);
                    float &c = coarsest[indices[row + x]];
                    c = std::fmax(c, cell);
                }
            }
        for (uint32_t s = 0; s < scene.source_count; ++s) {
            const float frequency = waveform_dominant_frequency(scene.sources[s].waveform);
            for (uint32_t m = 0; m < scene.material_count; ++m) {
                if (!(coarsest[m] > 0.0f)) continue;
                const float index = std::sqrt(scene.materials[m].relative_permittivity *
                                              scene.materials[m].relative_permeability);
                const float cells = kSpeedOfLight / (index * frequency * coarsest[m]);
                if (cells < 10.0f) {
                    std::snprintf(buffer, sizeof buffer,
                                  "%.1f cells per wavelength at %.3f GHz in index %.2f, where the "
                                  "cells are %.3f mm; below 10 the phase error is severe",
                                  static_cast<double>(cells), static_cast<double>(frequency * 1e-9f),
                                  static_cast<double>(index), static_cast<double>(coarsest[m] * 1e3f));
                    return buffer;
                }
            }
        }
    }
    static const char *names[3] = {"X", "Y", "Z"};
    for (uint32_t s = 0; s < scene.source_count; ++s) {
        const dp_source &src = scene.sources[s];
        if (src.shape != DP_SOURCE_PLANE_WAVE) continue;
        if (src.polarization == src.axis) {
            std::snprintf(buffer, sizeof buffer,
                          "a plane wave travelling along %s cannot be polarized along it; light is "
                          "transverse", names[src.axis % 3]);
            return buffer;
        }
        if (grid.axis[src.axis].worst_ratio() > 1.0f + 1e-6f) {
            std::snprintf(buffer, sizeof buffer,
                          "a plane wave needs a uniform grid along its travel axis, but %s is "
                          "graded; keep refinements off that axis", names[src.axis % 3]);
            return buffer;
        }
    }
    float size[3];
    for (int a = 0; a < 3; ++a) size[a] = grid.axis[a].length();
    for (uint32_t s = 0; s < scene.source_count; ++s) {
        const dp_source &src = scene.sources[s];
        float position[3] = {0.0f, 0.0f, 0.0f};
        if (src.shape == DP_SOURCE_POINT)
            for (int a = 0; a < 3; ++a) position[a] = src.at[a];
        else
            position[src.axis] = src.offset;
        if (!grid.contains(position)) {
            std::snprintf(buffer, sizeof buffer,
                          "source at [%g, %g, %g] m is outside the [%g, %g, %g] m domain; positions "
                          "are in metres from the centre, not cell indices",
                          position[0], position[1], position[2], size[0], size[1], size[2]);
            return buffer;
        }
    }
    for (uint32_t s = 0; s < scene.shape_count; ++s) {
        const dp_shape &sh = scene.shapes[s];
        float centre[3] = {0.0f, 0.0f, 0.0f}, half[3];
        if (sh.kind == DP_SHAPE_BLOCK) {
            for (int a = 0; a < 3; ++a) centre[a] = sh.center[a], half[a] = 0.5f * sh.size[a];
        } else if (sh.kind == DP_SHAPE_SPHERE) {
            for (int a = 0; a < 3; ++a) centre[a] = sh.center[a], half[a] = sh.radius;
        } else {
            for (int a = 0; a < 3; ++a) half[a] = INFINITY;
            centre[sh.axis] = sh.offset;
            half[sh.axis] = 0.5f * sh.thickness;
        }
        for (int a = 0; a < 3; ++a)
            if (std::fabs(centre[a]) - half[a] > 0.5f * size[a]) {
                std::snprintf(buffer, sizeof buffer,
                              "shape centred at [%g, %g, %g] m lies entirely outside the [%g, %g, "
                              "%g] m domain; positions are in metres from the centre",
                              centre[0], centre[1], centre[2], size[0], size[1], size[2]);
                return buffer;
            }
    }
    return {};
}

}  // namespace dpa
