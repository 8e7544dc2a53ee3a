// The device engine behind the C ABI: gpu::Simulation's counterpart
// (src/gpu.rs:159-799), built on HIP streams and hand-written gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "scene.hpp"

using namespace dpa;

#define DPA_VERSION "diaphane_amd 0.1.0 (gfx950)"

namespace {

struct HipFailure {
    hipError_t code;
    const char *what;
};

void check(hipError_t code, const char *what) {
    if (code != hipSuccess) throw HipFailure{code, what};
}

template <typename F>
int guarded(F &&body) {
    try {
        body();
        return DP_OK;
    } catch (const Error &e) {
        set_error(e.message);
        return e.code;
    } catch (const HipFailure &f) {
        set_error(std::string(f.what) + ": " + hipGetErrorString(f.code));
        return DP_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return DP_ERR_DEVICE;
    }
}

[[noreturn]] void fail(int code, const std::string &message) { throw Error{code, message}; }

// Device allocation owned by a sim.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    void allocate(size_t n) {
        release();
        if (n == 0) n = 16;
        check(hipMalloc(&ptr, n), "hipMalloc");
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(ptr);
    }
};

// Pinned host staging, grown on demand.
struct PinnedBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    ~PinnedBuffer() {
        if (ptr) (void)hipHostFree(ptr);
    }
    void reserve(size_t n) {
        if (n <= bytes) return;
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        check(hipHostMalloc(&ptr, n, hipHostMallocDefault), "hipHostMalloc");
        bytes = n;
    }
};

}  // namespace

struct dp_engine {
    int device = 0;
    std::string arch;
};

struct dp_sim {
    dp_engine *engine = nullptr;
    Grid grid;
    float time_step = 0.0f;
    std::vector<dp_source> sources;
    std::vector<int32_t> source_weight_offset;  // per source, −1 if unapodized
    uint32_t material_count = 0;
    bool reversible = false;
    uint64_t cells = 0;

    hipStream_t stream = nullptr;
    DeviceBuffer e[2], h[2];
    DeviceBuffer material, coefficients, axis_table, weights, eps_mu, widths;
    DeviceBuffer intensity, injections, partials;
    PinnedBuffer staging;
    int current = 0;  // which of the ping-pong pair holds the state

    uint64_t step = 0;
    bool accumulating = false;
    uint64_t accumulated = 0;
    int kernel = DP_KERNEL_FUSED;
    int rows = 0;  // 0: tuned default
    uint32_t depth = 0;  // 0: tuned default
    // Measured on MI355X at 768³: non-temporal stores help (~2%); renumbering
    // workgroups per XCD does not (the halos already hit L2).
    int tile_flags = DP_TILE_NONTEMPORAL;
    uint64_t launches = 0;
    hipEvent_t events[8] = {};

    ~dp_sim() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &ev : events)
            if (ev) (void)hipEventDestroy(ev);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void sync() { check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

    StepArgs args(int from, int to) const {
        StepArgs a{};
        a.e_in = e[from].as<float>();
        a.h_in = h[from].as<float>();
        a.e_out = e[to].as<float>();
        a.h_out = h[to].as<float>();
        a.material = material.as<uint8_t>();
        a.coefficients = coefficients.as<float4>();
        a.axis_table = axis_table.as<float>();
        a.weights = weights.as<float>();
        a.material_count = material_count;
        a.nx = grid.extent[0];
        a.ny = grid.extent[1];
        a.nz = grid.extent[2];
        a.cells = cells;
        return a;
    }

    int effective_rows() const {
        if (rows > 0) return rows;
        // Measured on MI355X at 768³: 4 rows (96 VGPRs) beats 6 (141) and 8 (178)
        // for the aligned kernel; the overlap kernel peaks at 6.
        return kernel == DP_KERNEL_FUSED_OVERLAP ? 6 : 4;
    }

    void ensure_intensity() {
        if (!intensity.ptr) {
            intensity.allocate(cells * sizeof(double));
            check(hipMemsetAsync(intensity.ptr, 0, cells * sizeof(double), stream), "hipMemsetAsync");
        }
    }

    void clear_intensity() {
        if (accumulated == 0) return;
        if (intensity.ptr)
            check(hipMemsetAsync(intensity.ptr, 0, cells * sizeof(double), stream), "hipMemsetAsync");
        accumulated = 0;
    }

    // Host-side injection tables for `steps` steps starting after `first`
    // (src/gpu.rs:442-465): per step the magnetic ones, then the electric
    // ones, each sign-scaled, zero values dropped like the reference does.
    struct StepInjections {
        uint32_t offset, magnetic, electric;
    };
    std::vector<StepInjections> build_injections(uint64_t first, uint64_t steps, int direction,
                                                 std::vector<InjDesc> &all) {
        std::vector<StepInjections> table(steps);
        all.clear();
        std::vector<InjDesc> electric;
        for (uint64_t s = 0; s < steps; ++s) {
            // forward step n → n+1 evaluates at (n+1)·Δt; reverse of step n at n·Δt
            const uint64_t n = direction > 0 ? first + s + 1 : first - s;
            const float time = static_cast<float>(n) * time_step;
            table[s].offset = static_cast<uint32_t>(all.size());
            electric.clear();
            for (size_t k = 0; k < sources.size(); ++k) {
                dp_injection pair[2];
                source_injections(sources[k], grid, time, pair);
                for (const dp_injection &q : pair) {
                    if (q.value == 0.0f) continue;
                    InjDesc d{};
                    std::memcpy(d.origin, q.origin, sizeof d.origin);
                    std::memcpy(d.extent, q.extent, sizeof d.extent);
                    d.component = q.component;
                    d.weight_offset = q.inverse_waist_squared == 0.0f ? -1 : source_weight_offset[k];
                    d.value = static_cast<float>(direction) * q.value;
                    (q.magnetic ? all : electric).push_back(d);
                }
            }
            table[s].magnetic = static_cast<uint32_t>(all.size()) - table[s].offset;
            table[s].electric = static_cast<uint32_t>(electric.size());
            all.insert(all.end(), electric.begin(), electric.end());
        }
        return table;
    }

    void upload_injections(const std::vector<InjDesc> &all) {
        const size_t bytes = std::max<size_t>(all.size(), 1) * sizeof(InjDesc);
        // The previous advance may still read the staging or device table.
        sync();
        staging.reserve(bytes);
        if (!all.empty()) std::memcpy(staging.ptr, all.data(), all.size() * sizeof(InjDesc));
        if (injections.bytes < bytes) injections.allocate(bytes * 2);
        check(hipMemcpyAsync(injections.ptr, staging.ptr, bytes, hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync");
    }

    void fill_tiles(StepArgs &a, int r, uint32_t owned_x) const {
        const uint32_t nx = grid.extent[0], ny = grid.extent[1], nz = grid.extent[2];
        a.tiles_x = (nx + owned_x - 1) / owned_x;
        a.tiles_y = (ny + (r - 2) - 1) / (r - 2);
        a.depth = std::min<uint32_t>(nz, depth ? depth : 32);
        a.tiles_z = (nz + a.depth - 1) / a.depth;
        a.tiles_total = a.tiles_x * a.tiles_y * a.tiles_z;
        a.xcd_remap = (tile_flags & DP_TILE_XCD_REMAP) ? 1 : 0;
    }

    void advance(uint64_t steps) {
        if (steps == 0) return;
        std::vector<InjDesc> all;
        auto table = build_injections(step, steps, +1, all);
        upload_injections(all);
        const InjDesc *device_inj = injections.as<InjDesc>();
        if (accumulating) ensure_intensity();
        const int r = effective_rows();
        for (uint64_t s = 0; s < steps; ++s) {
            const auto &t = table[s];
            if (kernel == DP_KERNEL_FUSED || kernel == DP_KERNEL_FUSED_OVERLAP) {
                StepArgs a = args(current, 1 - current);
                a.injections = device_inj + t.offset;
                a.n_magnetic = t.magnetic;
                a.n_electric = t.electric;
                a.intensity = accumulating ? intensity.as<double>() : nullptr;
                if (kernel == DP_KERNEL_FUSED) {
                    fill_tiles(a, r, kAlignedX);
                    launch_fused_aligned(a, r, (tile_flags & DP_TILE_NONTEMPORAL) != 0, stream);
                } else {
                    fill_tiles(a, r, kOwnedX);
                    launch_fused(a, r, stream);
                }
                current = 1 - current;
            } else {
                StepArgs a = args(current, current);
                launch_update_h(a, +1, stream);
                for (uint32_t q = 0; q < t.magnetic; ++q)
                    launch_inject(h[current].as<float>(), cells, grid.extent[0], grid.extent[1],
                                  all[t.offset + q], nullptr, weights.as<float>(), stream);
                launch_update_e(a, +1, stream);
                for (uint32_t q = 0; q < t.electric; ++q)
                    launch_inject(e[current].as<float>(), cells, grid.extent[0], grid.extent[1],
                                  all[t.offset + t.magnetic + q], nullptr, weights.as<float>(),
                                  stream);
                if (accumulating) launch_intensity(e[current].as<float>(), intensity.as<double>(), cells, stream);
            }
            check(hipGetLastError(), "kernel launch");
            ++launches;
            if (accumulating) ++accumulated;
        }
        step += steps;
    }

    void reverse(uint64_t steps) {
        if (steps == 0) return;
        if (!reversible)
            fail(DP_ERR_LOSSY,
                 "this scene is lossy, and running a dissipative update backwards amplifies "
                 "roundoff exponentially; reversal needs Boundary::Pec and no conductive material");
        if (step < steps) fail(DP_ERR_PAST_START, "cannot step back past the start");
        clear_intensity();
        std::vector<InjDesc> all;
        auto table = build_injections(step, steps, -1, all);
        upload_injections(all);
        for (uint64_t s = 0; s < steps; ++s) {
            const auto &t = table[s];
            StepArgs a = args(current, current);
            // Forward order is H, inject-H, E, inject-E; backward undoes each
            // in the opposite order (src/cpu.rs:266-272).
            for (uint32_t q = 0; q < t.electric; ++q)
                launch_inject(e[current].as<float>(), cells, grid.extent[0], grid.extent[1],
                              all[t.offset + t.magnetic + q], nullptr, weights.as<float>(), stream);
            launch_update_e(a, -1, stream);
            for (uint32_t q = 0; q < t.magnetic; ++q)
                launch_inject(h[current].as<float>(), cells, grid.extent[0], grid.extent[1],
                              all[t.offset + q], nullptr, weights.as<float>(), stream);
            launch_update_h(a, -1, stream);
            check(hipGetLastError(), "kernel launch");
        }
        step -= steps;
    }
};

namespace {

template <typename T>
void upload(DeviceBuffer &buffer, const std::vector<T> &host, hipStream_t stream) {
    buffer.allocate(std::max<size_t>(host.size(), 1) * sizeof(T));
    if (!host.empty())
        check(hipMemcpyAsync(buffer.ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void check_scene_pointers(const dp_scene *scene) {
    if (!scene) fail(DP_ERR_INVALID, "scene is null");
    if (scene->material_count && !scene->materials) fail(DP_ERR_INVALID, "materials pointer is null");
    if (scene->shape_count && !scene->shapes) fail(DP_ERR_INVALID, "shapes pointer is null");
    if (scene->source_count && !scene->sources) fail(DP_ERR_INVALID, "sources pointer is null");
}

}  // namespace

extern "C" {

const char *dp_last_error(void) { return last_error(); }
const char *dp_version(void) { return DPA_VERSION; }

int dp_grid_resolve(const dp_grid_spec *spec, uint32_t extent_out[3], float *widths_out,
                    size_t capacity, float *time_step_out) {
    return guarded([&] {
        if (!spec || !extent_out) fail(DP_ERR_INVALID, "null argument");
        Grid g = Grid::resolve(*spec);
        std::memcpy(extent_out, g.extent, sizeof g.extent);
        if (widths_out) {
            const size_t need = size_t(g.extent[0]) + g.extent[1] + g.extent[2];
            if (capacity < need) fail(DP_ERR_SIZE, "widths buffer is too small");
            size_t at = 0;
            for (int a = 0; a < 3; ++a)
                for (float w : g.axis[a].primary) widths_out[at++] = w;
        }
        if (time_step_out) *time_step_out = g.time_step();
    });
}

int dp_scene_validate(const dp_scene *scene) {
    return guarded([&] {
        check_scene_pointers(scene);
        std::string problem = validate(*scene);
        if (!problem.empty()) fail(DP_ERR_INVALID, problem);
    });
}

int dp_scene_material_indices(const dp_scene *scene, uint32_t *out, size_t count) {
    return guarded([&] {
        check_scene_pointers(scene);
        Grid g = Grid::resolve(scene->grid);
        if (count != g.cells()) fail(DP_ERR_SIZE, "index buffer length does not match the grid");
        rasterize(*scene, g, out);
    });
}

int dp_scene_tables(const dp_scene *scene, float *coefficients, size_t coefficient_capacity,
                    float *axis_table, size_t axis_capacity) {
    return guarded([&] {
        check_scene_pointers(scene);
        Grid g = Grid::resolve(scene->grid);
        if (coefficients) {
            if (coefficient_capacity < 4 * size_t(scene->material_count))
                fail(DP_ERR_SIZE, "coefficient buffer too small");
            const float dt = g.time_step();
            for (uint32_t m = 0; m < scene->material_count; ++m) {
                Coefficients c = material_coefficients(scene->materials[m], dt);
                coefficients[4 * m + 0] = c.electric_gain;
                coefficients[4 * m + 1] = c.electric_loss;
                coefficients[4 * m + 2] = c.magnetic_gain;
                coefficients[4 * m + 3] = c.magnetic_loss;
            }
        }
        if (axis_table) {
            const size_t span = size_t(g.extent[0]) + g.extent[1] + g.extent[2];
            if (axis_capacity < 4 * span) fail(DP_ERR_SIZE, "axis table buffer too small");
            Absorber absorber = Absorber::build(g, *scene);
            size_t at = 0;
            for (int a = 0; a < 3; at += g.extent[a], ++a) {
                const auto gh = g.gains(g.axis[a].primary), ge = g.gains(g.axis[a].dual);
                for (uint32_t i = 0; i < g.extent[a]; ++i) {
                    axis_table[kGainH * span + at + i] = gh[i];
                    axis_table[kGainE * span + at + i] = ge[i];
                    axis_table[kLossInt * span + at + i] = absorber.integer[a][i];
                    axis_table[kLossHalf * span + at + i] = absorber.half[a][i];
                }
            }
        }
    });
}

int dp_scene_injections(const dp_scene *scene, float time, dp_injection *out, size_t capacity) {
    return guarded([&] {
        check_scene_pointers(scene);
        if (capacity < 2 * size_t(scene->source_count)) fail(DP_ERR_SIZE, "injection buffer too small");
        Grid g = Grid::resolve(scene->grid);
        for (uint32_t s = 0; s < scene->source_count; ++s)
            source_injections(scene->sources[s], g, time, out + 2 * s);
    });
}

int dp_engine_create(int device, dp_engine **out) {
    return guarded([&] {
        if (!out) fail(DP_ERR_INVALID, "null argument");
        int count = 0;
        check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) fail(DP_ERR_DEVICE, "no usable GPU device");
        check(hipSetDevice(device), "hipSetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        auto engine = std::make_unique<dp_engine>();
        engine->device = device;
        engine->arch = prop.gcnArchName;
        *out = engine.release();
    });
}

void dp_engine_destroy(dp_engine *engine) { delete engine; }

int dp_engine_arch(dp_engine *engine, char *out, size_t capacity) {
    return guarded([&] {
        if (!engine || !out || capacity == 0) fail(DP_ERR_INVALID, "null argument");
        std::snprintf(out, capacity, "%s", engine->arch.c_str());
    });
}

int dp_sim_create(dp_engine *engine, const dp_scene *scene, dp_sim **out) {
    return guarded([&] {
        if (!engine || !out) fail(DP_ERR_INVALID, "null argument");
        check_scene_pointers(scene);
        check(hipSetDevice(engine->device), "hipSetDevice");
        auto sim = std::make_unique<dp_sim>();
        sim->engine = engine;
        sim->grid = Grid::resolve(scene->grid);  // validates (src/gpu.rs:191)
        const Grid &g = sim->grid;
        if (scene->material_count == 0) fail(DP_ERR_INVALID, "the material table is empty; slot 0 must be vacuum");
        if (scene->material_count > 256)
            fail(DP_ERR_UNSUPPORTED, "this build stores material indices as u8: at most 256 materials");
        for (uint32_t s = 0; s < scene->shape_count; ++s)
            if (scene->shapes[s].material >= scene->material_count)
                fail(DP_ERR_INVALID, "shape references a material past the end of the table");
        for (uint32_t s = 0; s < scene->source_count; ++s) {
            const dp_source &src = scene->sources[s];
            if (src.shape > DP_SOURCE_PLANE_WAVE || src.polarization > 2 || src.axis > 2)
                fail(DP_ERR_INVALID, "source has an unknown shape or axis");
        }
        sim->cells = g.cells();
        sim->time_step = g.time_step();
        sim->material_count = scene->material_count;
        sim->sources.assign(scene->sources, scene->sources + scene->source_count);
        Absorber absorber = Absorber::build(g, *scene);

        // is_reversible (src/cpu.rs:203-209): no loss anywhere.
        std::vector<float4> coeff(scene->material_count);
        std::vector<float2> eps_mu(scene->material_count);
        bool lossless = true;
        for (uint32_t m = 0; m < scene->material_count; ++m) {
            Coefficients c = material_coefficients(scene->materials[m], sim->time_step);
            coeff[m] = make_float4(c.electric_gain, c.electric_loss, c.magnetic_gain, c.magnetic_loss);
            eps_mu[m] = make_float2(scene->materials[m].relative_permittivity,
                                    scene->materials[m].relative_permeability);
            lossless = lossless && c.electric_loss == 0.0f && c.magnetic_loss == 0.0f;
        }
        sim->reversible = lossless && absorber.peak == 0.0f;

        // Per-axis table: gains (primary, dual) and absorber (integer, half).
        const uint32_t span = g.extent[0] + g.extent[1] + g.extent[2];
        std::vector<float> table(4 * span), widths(2 * span);
        for (int a = 0, at = 0; a < 3; at += g.extent[a], ++a) {
            const auto gh = g.gains(g.axis[a].primary), ge = g.gains(g.axis[a].dual);
            for (uint32_t i = 0; i < g.extent[a]; ++i) {
                table[kGainH * span + at + i] = gh[i];
                table[kGainE * span + at + i] = ge[i];
                table[kLossInt * span + at + i] = absorber.integer[a][i];
                table[kLossHalf * span + at + i] = absorber.half[a][i];
                widths[at + i] = g.axis[a].primary[i];
                widths[span + at + i] = g.axis[a].dual[i];
            }
        }

        // Apodization weights per sheet source over its (fixed) box.
        std::vector<float> weights;
        sim->source_weight_offset.assign(sim->sources.size(), -1);
        for (size_t s = 0; s < sim->sources.size(); ++s) {
            dp_injection pair[2];
            source_injections(sim->sources[s], g, 0.0f, pair);
            const dp_injection &q = pair[0];
            if (q.inverse_waist_squared == 0.0f) continue;
            sim->source_weight_offset[s] = static_cast<int32_t>(weights.size());
            for (uint32_t dz = 0; dz < q.extent[2]; ++dz)
                for (uint32_t dy = 0; dy < q.extent[1]; ++dy)
                    for (uint32_t dx = 0; dx < q.extent[0]; ++dx) {
                        const uint32_t c[3] = {q.origin[0] + dx, q.origin[1] + dy, q.origin[2] + dz};
                        weights.push_back(injection_weight(q, g, c));
                    }
        }
        if (weights.size() > size_t(INT32_MAX)) fail(DP_ERR_UNSUPPORTED, "apodization table too large");

        std::vector<uint8_t> indices(sim->cells);
        rasterize_u8(*scene, g, indices.data());

        check(hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto &ev : sim->events) check(hipEventCreate(&ev), "hipEventCreate");
        const size_t field_bytes = 3 * sim->cells * sizeof(float);
        for (int b = 0; b < 2; ++b) {
            sim->e[b].allocate(field_bytes);
            sim->h[b].allocate(field_bytes);
        }
        check(hipMemsetAsync(sim->e[0].ptr, 0, field_bytes, sim->stream), "hipMemsetAsync");
        check(hipMemsetAsync(sim->h[0].ptr, 0, field_bytes, sim->stream), "hipMemsetAsync");
        upload(sim->material, indices, sim->stream);
        upload(sim->coefficients, coeff, sim->stream);
        upload(sim->eps_mu, eps_mu, sim->stream);
        upload(sim->axis_table, table, sim->stream);
        upload(sim->widths, widths, sim->stream);
        upload(sim->weights, weights, sim->stream);
        sim->partials.allocate(2 * sizeof(double) * energy_blocks(g.extent[0], g.extent[1], g.extent[2]));
        sim->injections.allocate(64 * sizeof(InjDesc));
        sim->sync();
        *out = sim.release();
    });
}

void dp_sim_destroy(dp_sim *sim) { delete sim; }

int dp_sim_extent(const dp_sim *sim, uint32_t out[3]) {
    if (!sim || !out) return DP_ERR_INVALID;
    std::memcpy(out, sim->grid.extent, sizeof sim->grid.extent);
    return DP_OK;
}

uint64_t dp_sim_step_count(const dp_sim *sim) { return sim ? sim->step : 0; }
float dp_sim_time(const dp_sim *sim) {
    return sim ? static_cast<float>(sim->step) * sim->time_step : 0.0f;
}
float dp_sim_time_step(const dp_sim *sim) { return sim ? sim->time_step : 0.0f; }
uint64_t dp_sim_field_bytes(const dp_sim *sim) { return sim ? 3 * sim->cells * sizeof(float) : 0; }
uint64_t dp_sim_bytes_per_step(const dp_sim *sim) { return sim ? 4 * dp_sim_field_bytes(sim) : 0; }
int dp_sim_is_reversible(const dp_sim *sim) { return sim && sim->reversible ? 1 : 0; }
uint64_t dp_sim_launch_count(const dp_sim *sim) { return sim ? sim->launches : 0; }

int dp_sim_advance(dp_sim *sim, uint64_t steps) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        sim->advance(steps);
    });
}

int dp_sim_reverse(dp_sim *sim, uint64_t steps) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        sim->reverse(steps);
    });
}

int dp_sim_wait(dp_sim *sim, uint32_t timeout_ms) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        const uint32_t limit = timeout_ms ? timeout_ms : 120000;  // src/gpu.rs:184
        const auto start = std::chrono::steady_clock::now();
        for (;;) {
            hipError_t status = hipStreamQuery(sim->stream);
            if (status == hipSuccess) return;
            if (status != hipErrorNotReady) check(status, "hipStreamQuery");
            if (std::chrono::steady_clock::now() - start > std::chrono::milliseconds(limit))
                fail(DP_ERR_TIMEOUT, "the GPU did not finish in time");
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    });
}

int dp_sim_reset(dp_sim *sim) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        const size_t field_bytes = 3 * sim->cells * sizeof(float);
        check(hipMemsetAsync(sim->e[sim->current].ptr, 0, field_bytes, sim->stream), "hipMemsetAsync");
        check(hipMemsetAsync(sim->h[sim->current].ptr, 0, field_bytes, sim->stream), "hipMemsetAsync");
        if (sim->intensity.ptr)
            check(hipMemsetAsync(sim->intensity.ptr, 0, sim->cells * sizeof(double), sim->stream),
                  "hipMemsetAsync");
        sim->step = 0;
        sim->accumulated = 0;
    });
}

int dp_sim_set_accumulate_intensity(dp_sim *sim, int on) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        if (on && !sim->accumulating) sim->clear_intensity();
        if (on) sim->ensure_intensity();
        sim->accumulating = on != 0;
    });
}

uint64_t dp_sim_accumulated_steps(const dp_sim *sim) { return sim ? sim->accumulated : 0; }

static int read_field(dp_sim *sim, const DeviceBuffer *pair, float *out, size_t count) {
    return guarded([&] {
        if (!sim || !out) fail(DP_ERR_INVALID, "null argument");
        if (count != 3 * sim->cells) fail(DP_ERR_SIZE, "buffer length must be 3 x cell count");
        sim->sync();
        check(hipMemcpy(out, pair[sim->current].ptr, count * sizeof(float), hipMemcpyDeviceToHost),
              "hipMemcpy");
    });
}

int dp_sim_read_electric(dp_sim *sim, float *out, size_t count) {
    return read_field(sim, sim ? sim->e : nullptr, out, count);
}

int dp_sim_read_magnetic(dp_sim *sim, float *out, size_t count) {
    return read_field(sim, sim ? sim->h : nullptr, out, count);
}

int dp_sim_read_intensity_f64(dp_sim *sim, double *out, size_t count) {
    return guarded([&] {
        if (!sim || !out) fail(DP_ERR_INVALID, "null argument");
        if (count != sim->cells) fail(DP_ERR_SIZE, "buffer length must be the cell count");
        sim->sync();
        if (!sim->intensity.ptr) {
            std::fill(out, out + count, 0.0);
            return;
        }
        check(hipMemcpy(out, sim->intensity.ptr, count * sizeof(double), hipMemcpyDeviceToHost),
              "hipMemcpy");
    });
}

int dp_sim_read_intensity(dp_sim *sim, float *out, size_t count) {
    return guarded([&] {
        if (!sim || !out) fail(DP_ERR_INVALID, "null argument");
        std::vector<double> wide(count);
        int status = dp_sim_read_intensity_f64(sim, wide.data(), count);
        if (status != DP_OK) throw Error{status, last_error()};
        for (size_t i = 0; i < count; ++i) out[i] = static_cast<float>(wide[i]);
    });
}

int dp_sim_write_fields(dp_sim *sim, const float *electric, const float *magnetic, size_t count,
                        uint64_t step) {
    return guarded([&] {
        if (!sim || !electric || !magnetic) fail(DP_ERR_INVALID, "null argument");
        sim->clear_intensity();  // the average belongs to the abandoned history
        if (count != 3 * sim->cells) fail(DP_ERR_SIZE, "snapshot was taken from a different grid");
        sim->sync();
        check(hipMemcpy(sim->e[sim->current].ptr, electric, count * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy");
        check(hipMemcpy(sim->h[sim->current].ptr, magnetic, count * sizeof(float), hipMemcpyHostToDevice),
              "hipMemcpy");
        sim->step = step;
    });
}

int dp_sim_energy(dp_sim *sim, double *electric, double *magnetic) {
    return guarded([&] {
        if (!sim || !electric || !magnetic) fail(DP_ERR_INVALID, "null argument");
        const uint32_t *n = sim->grid.extent;
        const uint32_t blocks =
            launch_energy(sim->e[sim->current].as<float>(), sim->h[sim->current].as<float>(),
                          sim->material.as<uint8_t>(), sim->eps_mu.as<float2>(),
                          sim->widths.as<float>(), n[0], n[1], n[2], sim->partials.as<double>(),
                          sim->stream);
        check(hipGetLastError(), "kernel launch");
        std::vector<double> partials(2 * blocks);
        sim->sync();
        check(hipMemcpy(partials.data(), sim->partials.ptr, partials.size() * sizeof(double),
                        hipMemcpyDeviceToHost),
              "hipMemcpy");
        double ue = 0.0, uh = 0.0;
        for (uint32_t b = 0; b < blocks; ++b) {
            ue += partials[2 * b];
            uh += partials[2 * b + 1];
        }
        *electric = ue;
        *magnetic = uh;
    });
}

int dp_sim_device_ptr(dp_sim *sim, int which, void **out) {
    return guarded([&] {
        if (!sim || !out) fail(DP_ERR_INVALID, "null argument");
        switch (which) {
        case DP_BUFFER_ELECTRIC: *out = sim->e[sim->current].ptr; break;
        case DP_BUFFER_MAGNETIC: *out = sim->h[sim->current].ptr; break;
        case DP_BUFFER_INTENSITY: sim->ensure_intensity(); *out = sim->intensity.ptr; break;
        case DP_BUFFER_MATERIAL_INDEX: *out = sim->material.ptr; break;
        default: fail(DP_ERR_INVALID, "unknown buffer");
        }
    });
}

int dp_sim_set_kernel(dp_sim *sim, int kernel) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        if (kernel != DP_KERNEL_FUSED && kernel != DP_KERNEL_TWO_PASS && kernel != DP_KERNEL_FUSED_OVERLAP)
            fail(DP_ERR_INVALID, "unknown kernel");
        sim->kernel = kernel;
    });
}

int dp_sim_set_tile(dp_sim *sim, int rows, int depth, int flags) {
    return guarded([&] {
        if (!sim) fail(DP_ERR_INVALID, "null sim");
        static const int allowed[] = {0, 4, 6, 8, 10, 12};
        if (std::find(std::begin(allowed), std::end(allowed), rows) == std::end(allowed))
            fail(DP_ERR_INVALID, "rows must be one of 4, 6, 8, 10, 12 (or 0 for the default)");
        if (depth < 0) fail(DP_ERR_INVALID, "depth must be non-negative");
        sim->rows = rows;
        sim->depth = static_cast<uint32_t>(depth);
        if (flags >= 0) sim->tile_flags = flags & (DP_TILE_XCD_REMAP | DP_TILE_NONTEMPORAL);
    });
}

int dp_sim_stream_probe(dp_sim *sim, int iterations, double *gbs) {
    return guarded([&] {
        if (!sim || !gbs || iterations <= 0) fail(DP_ERR_INVALID, "bad argument");
        const int from = sim->current, to = 1 - sim->current;
        auto launch = [&] {
            launch_stream_probe(sim->e[from].as<float>(), sim->h[from].as<float>(), sim->material.as<uint8_t>(),
                                sim->e[to].as<float>(), sim->h[to].as<float>(), sim->cells, sim->stream);
        };
        launch();
        check(hipGetLastError(), "kernel launch");
        check(hipEventRecord(sim->events[6], sim->stream), "hipEventRecord");
        for (int i = 0; i < iterations; ++i) launch();
        check(hipEventRecord(sim->events[7], sim->stream), "hipEventRecord");
        check(hipEventSynchronize(sim->events[7]), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, sim->events[6], sim->events[7]), "hipEventElapsedTime");
        const double bytes = 49.0 * double(sim->cells / 4 * 4) * iterations;
        *gbs = bytes / (double(ms) / 1e3) / 1e9;
    });
}

int dp_sim_record_event(dp_sim *sim, int slot) {
    return guarded([&] {
        if (!sim || slot < 0 || slot >= 8) fail(DP_ERR_INVALID, "bad event slot");
        check(hipEventRecord(sim->events[slot], sim->stream), "hipEventRecord");
    });
}

int dp_sim_elapsed_ms(dp_sim *sim, int from_slot, int to_slot, float *ms) {
    return guarded([&] {
        if (!sim || !ms || from_slot < 0 || from_slot >= 8 || to_slot < 0 || to_slot >= 8)
            fail(DP_ERR_INVALID, "bad event slot");
        check(hipEventSynchronize(sim->events[to_slot]), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, sim->events[from_slot], sim->events[to_slot]),
              "hipEventElapsedTime");
    });
}

}  // extern "C"
