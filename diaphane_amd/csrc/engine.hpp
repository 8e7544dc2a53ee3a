// Device engine internals shared by the single-sim and batch entry points:
// the sim object (gpu::Simulation's counterpart, src/gpu.rs:159-799) and
// the RAII helpers around HIP allocations.
#pragma once

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


namespace dpa_detail {

struct HipFailure {
    hipError_t code;
    const char *what;
};

inline void check(hipError_t code, const char *what) {
    if (code != hipSuccess) throw HipFailure{code, what};
}

template <typename F>
inline int guarded(F &&body) {
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

[[noreturn]] inline void fail(int code, const std::string &message) { throw Error{code, message}; }

// A copy ordered on `stream`, waited for. The library never uses the legacy
// NULL stream: HIP refuses a legacy-stream copy while any stream of the
// device is capturing a graph (a batch on another thread may be).
inline void copy_on(hipStream_t stream, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    check(hipMemcpyAsync(dst, src, bytes, kind, stream), "hipMemcpyAsync");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

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
    // Grows to at least n bytes (twice n when it must grow); the contents
    // are not kept.
    void reserve(size_t n) {
        if (n > bytes) allocate(2 * n);
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

// Pinned host staging, grown on demand with headroom (twice the request),
// so a run of slowly growing requests re-allocates only a few times. The
// caller makes sure no pending copy reads the old block before it grows.
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
        bytes = 0;
        check(hipHostMalloc(&ptr, 2 * n, hipHostMallocDefault), "hipHostMalloc");
        bytes = 2 * n;
    }
};

}  // namespace dpa_detail

using namespace dpa_detail;

namespace dpa {
// Takes the device address of the fixed scratch array that
// dp_sim_measure_peak_noalloc uses, once per device (noalloc.hip); called
// by dp_engine_create with `device` current.
void prepare_noalloc_scratch(int device);
}  // namespace dpa

struct dp_batch;

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
    DeviceBuffer intensity, partials;
    DeviceBuffer empty_injections;  // one entry, for launches with no injections
    DeviceBuffer peak_partials;  // per-block maxima of dp_sim_measure_peak
    DeviceBuffer intensity_f32;  // DP_BUFFER_INTENSITY_F32, refreshed on request
    DeviceBuffer lookup;         // DP_BUFFER_LOOKUP: Grid::cell_lookup, 3 × lookup_samples
    uint32_t lookup_samples = 0;
    // Injection tables of the advances in flight. Each advance fills the next
    // slot's pinned staging, copies it to the slot's device table and records
    // `done` after its last launch, so an advance waits only for the one that
    // used its slot kInjectionSlots advances earlier (long finished, unless
    // the host runs that far ahead), never for the previous one.
    struct InjectionSlot {
        PinnedBuffer staging;
        DeviceBuffer table;
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    static constexpr int kInjectionSlots = 4;
    InjectionSlot injection_slots[kInjectionSlots];
    int next_injection_slot = 0;
    // One step's injections in a table: magnetic ones from `offset`, then
    // the electric ones.
    struct StepInjections {
        uint32_t offset, magnetic, electric;
    };
    // The forward injection tables of steps first + 1 to first + steps,
    // uploaded once to a slot and reused by every advance they cover: a
    // step's injections depend only on its number (the sources are fixed at
    // creation). An advance of a few steps, as a viewer makes each frame,
    // otherwise built its table on the host and copied it to the device on
    // every call. A window is never written after its upload; the next one
    // takes another slot. Its slot's event is recorded once, when the window
    // is retired, after every launch queued with it (one record per call put
    // a marker between the launches of one-step calls: 15.5 against 12.9
    // µs a step at 48³).
    struct InjectionWindow {
        bool valid = false;
        uint64_t first = 0, steps = 0;
        std::vector<StepInjections> table;
        std::vector<InjDesc> all;         // host copy (the two-pass path launches from it)
        const InjDesc *device = nullptr;  // its device table
        InjectionSlot *slot = nullptr;    // the slot holding it (nullptr: no injections)
    } window;
    static constexpr uint64_t kInjectionWindowSteps = 64;
    int current = 0;  // which of the ping-pong pair holds the state

    uint64_t step = 0;
    bool accumulating = false;
    uint64_t accumulated = 0;
    int kernel = DP_KERNEL_FUSED;
    // Fused-kernel tiling for plain steps [0] and for steps that accumulate
    // intensity [1], whose extra read-modify-write can favour another shape.
    // rows/depth 0 select the defaults. Measured on MI355X at 768³:
    // non-temporal stores help (~2%); renumbering workgroups per XCD does not.
    struct Tiling {
        int rows = 0;
        uint32_t depth = 0;
        int flags = DP_TILE_NONTEMPORAL;
    };
    // The one-step shape of `t` (itself unless it takes two steps a launch):
    // with DP_TILE_TWO_STEP, the launch that takes an odd count's last step
    // runs 4 rows with the size rule's depth and `t`'s store, addressing and
    // XCD flags. The two-step rows do not carry over: 8 rows a wave ran the
    // one-step kernel at 80³ 3× slower than 4 with a plane-wave source (93
    // against 29 µs). The autotuner's fastest one-step shape, taken from its
    // one short sweep, was not kept for it: its pick varied from run to run
    // and once cost the 80³ readback job 40%.
    static Tiling one_step_tiling(const Tiling &t) {
        if (!(t.flags & DP_TILE_TWO_STEP)) return t;
        Tiling one;
        one.rows = 4;
        one.depth = 0;
        one.flags = t.flags & (DP_TILE_NONTEMPORAL | DP_TILE_BUFFER | DP_TILE_XCD_REMAP | 0xFF00);
        return one;
    }
    Tiling tiling[2];
    // Grids from this size start on stacked two-step tiles; below it one step
    // per launch (4 rows) is faster (measured on MI355X: 64³ 11.0 against
    // 14.6 µs a step, 96³ 26.8 against 14.2 µs, 160³ 66 against 42 µs).
    static constexpr uint64_t kTwoStepDefaultCells = 768ull << 10;
    // Two steps per launch as the size rule and the autotuner set it: with
    // XCD-grouped workgroups (DP_TILE_XCD_REMAP, runs of 8).
    static constexpr int kTwoStepTileFlags = DP_TILE_TWO_STEP | DP_TILE_XCD_REMAP;
    // The large-grid default: two steps per launch with four waves of 4
    // rows stacked along y (16-row tiles, 12 owned), XCD runs of 16
    // workgroups. Measured on MI355X at 768³: 3.18 ms per plain step and
    // 4.38 ms per intensity step, against 3.34 and 4.75 ms with one 10-row
    // wave per tile; runs of 16 against 8 at 12-plane slabs, 6.15 / 7.90
    // against 6.15–6.19 / 7.93–7.96 ms per plain / intensity launch (runs of
    // 2, 4 and 32 were slower).
    static constexpr int kStackTileFlags = kTwoStepTileFlags | DP_TILE_STACK | (4 << 16) | (16 << 8);
    const Tiling &active() const { return tiling[accumulating ? 1 : 0]; }
    uint64_t launches = 0;
    hipEvent_t events[8] = {};

    ~dp_sim();  // leaves its batch first (batch.cpp)

    // Overlapped readbacks (dp_sim_begin_read): the fields are copied to a
    // slot's `stage` on the sim's stream, then to the caller's host buffer on
    // `read_stream`, which waits for the slot's `staged`; `done` follows the
    // host copy. Two slots, used in turn, so a readback can be queued while
    // the previous one is still on its way to the host; the host copies run
    // in order on the one stream.
    struct ReadSlot {
        DeviceBuffer stage;
        hipEvent_t staged = nullptr, done = nullptr;
        bool pending = false;
    };
    static constexpr int kReadSlots = 2;
    static_assert(kReadSlots == 2, "wait_reads picks the older of two slots");
    ReadSlot read_slots[kReadSlots];
    int next_read_slot = 0;  // the slot the next readback takes
    hipStream_t read_stream = nullptr;

    int reads_in_flight() const {
        int n = 0;
        for (const ReadSlot &r : read_slots) n += r.pending ? 1 : 0;
        return n;
    }

    // Waits until at most `keep` readbacks are in flight, the oldest first
    // (timeout 0: 120 s).
    void wait_reads(int keep, uint32_t timeout_ms) {
        const auto limit = std::chrono::milliseconds(timeout_ms ? timeout_ms : 120000);
        const auto start = std::chrono::steady_clock::now();
        while (reads_in_flight() > keep) {
            // the older of two in flight is the slot the next readback takes
            ReadSlot &r = read_slots[read_slots[next_read_slot].pending ? next_read_slot : 1 - next_read_slot];
            for (;;) {
                const hipError_t status = hipEventQuery(r.done);
                if (status == hipSuccess) break;
                if (status != hipErrorNotReady) check(status, "hipEventQuery");
                const auto waited = std::chrono::steady_clock::now() - start;
                if (waited > limit) fail(DP_ERR_TIMEOUT, "the readback did not finish in time");
                // Polled without sleeping for the first 2 ms: a 50 µs sleep
                // woke 60–100 µs late, and a caller that queues the next
                // steps after this returns left the device idle meanwhile.
                if (waited < std::chrono::milliseconds(2))
                    std::this_thread::yield();
                else
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
            r.pending = false;
        }
    }

    void release_stream() {
        for (ReadSlot &r : read_slots) {
            if (r.pending) (void)hipEventSynchronize(r.done);
            r.pending = false;
            if (r.staged) (void)hipEventDestroy(r.staged);
            if (r.done) (void)hipEventDestroy(r.done);
        }
        if (read_stream) (void)hipStreamDestroy(read_stream);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &ev : events)
            if (ev) (void)hipEventDestroy(ev);
        for (auto &slot : injection_slots)
            if (slot.done) (void)hipEventDestroy(slot.done);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Set while a batch advances this sim. A member's `stream` is the batch's
    // stream, so its own copies, clears and waits are ordered with the batch
    // launches; `own_stream` keeps the sim's stream until it leaves.
    dp_batch *batch = nullptr;
    hipStream_t own_stream = nullptr;

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
        if (active().rows > 0) return active().rows;
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

    // Fills the next injection slot with `all` and queues its copy to the
    // device; the caller launches from the slot's table and then calls
    // release_injections. An empty list takes no slot (nullptr): it copies
    // nothing and launches from the one-entry `empty_injections`, which no
    // kernel reads with zero counts.
    InjectionSlot *upload_injections(const std::vector<InjDesc> &all, const InjDesc *&device_table) {
        if (all.empty()) {
            device_table = empty_injections.as<InjDesc>();
            return nullptr;
        }
        InjectionSlot &slot = injection_slots[next_injection_slot];
        next_injection_slot = (next_injection_slot + 1) % kInjectionSlots;
        if (window.valid && window.slot == &slot) retire_window();  // the ring came round to it
        if (slot.pending) check(hipEventSynchronize(slot.done), "hipEventSynchronize");
        slot.pending = false;
        const size_t bytes = all.size() * sizeof(InjDesc);
        slot.staging.reserve(bytes);
        std::memcpy(slot.staging.ptr, all.data(), bytes);
        if (slot.table.bytes < bytes) slot.table.allocate(bytes * 2);
        check(hipMemcpyAsync(slot.table.ptr, slot.staging.ptr, bytes, hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync");
        device_table = slot.table.as<InjDesc>();
        return &slot;
    }

    // Ends the injection window: its slot stays busy until the stream passes
    // the launches queued with it.
    void retire_window() {
        if (!window.valid) return;
        window.valid = false;
        release_injections(window.slot);
        window.slot = nullptr;
    }

    // Marks the slot busy until the stream passes this point (after the
    // launches that read its table). Should the record fail, the stream is
    // drained instead, so the slot is free for its next use either way.
    void release_injections(InjectionSlot *slot) {
        if (!slot) return;
        if (hipEventRecord(slot->done, stream) == hipSuccess)
            slot->pending = true;
        else
            (void)hipStreamSynchronize(stream);
    }

    // The default z-slab depth for a grid whose x-y plane holds `tiles_xy`
    // tiles: a wave marches its slab plane after plane, so a grid with few
    // tiles per plane needs short slabs to put enough waves on the device,
    // and a large one long slabs to amortize each slab's halo plane. The
    // power of two at or below tiles_xy / 50, within 1..32, is within about
    // 10% of the best depth measured on MI355X from 32³ to 768³ (32³: 1,
    // 160³: 4, 256³: 8, 768³: 32), where a fixed 32 is up to 9× slower on
    // small grids.
    static uint32_t default_depth(uint32_t tiles_xy) {
        uint32_t depth = 1;
        while (depth < 32 && depth * 2 * 50 <= tiles_xy) depth *= 2;
        return depth;
    }
    // Two steps per launch re-read two halo planes per slab instead of one,
    // so they favour slabs about twice as deep: the power of two at or below
    // tiles_xy / 25, within 2..32 (measured best on MI355X: 16 at 384³ and
    // 512³, 32 at 768³).
    static uint32_t default_pair_depth(uint32_t tiles_xy) {
        uint32_t depth = 2;
        while (depth < 32 && depth * 2 * 25 <= tiles_xy) depth *= 2;
        return depth;
    }

    // `halo_rows`: register rows a wave does not own (2 for one step per
    // launch, 4 for two).
    // `waves_per_tile` (stacked tiles) counts in the default depth, which is
    // sized to put enough waves on the device.
    void fill_tiles(StepArgs &a, int r, uint32_t owned_x, const Tiling &t, int halo_rows = 2,
                    uint32_t waves_per_tile = 1) const {
        const uint32_t nx = grid.extent[0], ny = grid.extent[1], nz = grid.extent[2];
        const uint32_t owned_y = uint32_t(r - halo_rows);
        // stacked two-step tiles own more rows at the walls
        a.tiles_x = (nx + owned_x - 1) / owned_x;
        a.tiles_y = waves_per_tile > 1 ? stack_tiles_y(ny, uint32_t(r)) : (ny + owned_y - 1) / owned_y;
        const uint32_t tiles_xy = a.tiles_x * a.tiles_y;
        const uint32_t waves_xy = tiles_xy * waves_per_tile;
        a.depth = std::min<uint32_t>(
            nz, t.depth ? t.depth : (halo_rows == 4 ? default_pair_depth(waves_xy) : default_depth(waves_xy)));
        a.tiles_z = (nz + a.depth - 1) / a.depth;
        a.tiles_total = a.tiles_x * a.tiles_y * a.tiles_z;
        if (t.flags & DP_TILE_XCD_REMAP) {
            const uint32_t group = (uint32_t(t.flags) >> 8) & 0xFFu;
            a.xcd_remap = group ? group : 8;
        } else {
            a.xcd_remap = 0;
        }
    }

    // Register rows of the two-step kernel for a tiling (6, 8 or 10).
    static int pair_rows(const Tiling &t) { return t.rows == 6 || t.rows == 10 ? t.rows : 8; }
    // DP_TILE_STACK: register rows per wave (4 or 6) and waves per tile (2 or
    // 4, from bits 16..19 of the flags).
    static bool stacked(const Tiling &t) { return (t.flags & DP_TILE_STACK) && (t.flags & DP_TILE_TWO_STEP); }
    static int stack_rows(const Tiling &t) { return t.rows == 4 ? 4 : 6; }
    static int stack_waves(int flags) { return ((flags >> 16) & 0xF) == 2 ? 2 : 4; }

    // Cuts `a` into the two-step tiles of `tl` and launches the kernel that
    // tiling names: one wave per tile, or waves stacked along y.
    void launch_pair_args(StepArgs &a, const Tiling &tl, hipStream_t on) const {
        const bool buffer = (tl.flags & DP_TILE_BUFFER) != 0;
        if (stacked(tl)) {
            const int r = stack_rows(tl), w = stack_waves(tl.flags);
            fill_tiles(a, w * r, kPairOwnedX, tl, 4, uint32_t(w));
            launch_fused_stack(a, r, w, buffer, on);
            return;
        }
        const int pr = pair_rows(tl);
        fill_tiles(a, pr, kPairOwnedX, tl, 4);
        launch_fused_pair(a, pr, (tl.flags & DP_TILE_NONTEMPORAL) != 0, buffer, on);
    }

    // One launch of the two-step kernel: steps `first` and `first + 1` of
    // `table`.
    void launch_pair(const StepInjections *table, uint64_t first, const InjDesc *device_inj) {
        const Tiling &tl = active();
        StepArgs a = args(current, 1 - current);
        a.injections = device_inj + table[first].offset;
        a.n_magnetic = table[first].magnetic;
        a.n_electric = table[first].electric;
        a.injections2 = device_inj + table[first + 1].offset;
        a.n_magnetic2 = table[first + 1].magnetic;
        a.n_electric2 = table[first + 1].electric;
        a.intensity = accumulating ? intensity.as<double>() : nullptr;
        launch_pair_args(a, tl, stream);
        current = 1 - current;
    }

    void advance(uint64_t steps) {
        if (steps == 0) return;
        if (!window.valid || step < window.first || step + steps > window.first + window.steps) {
            retire_window();
            const uint64_t n = std::max<uint64_t>(steps, kInjectionWindowSteps);
            window.table = build_injections(step, n, +1, window.all);
            window.slot = upload_injections(window.all, window.device);
            window.first = step;
            window.steps = n;
            window.valid = true;
        }
        const StepInjections *table = window.table.data() + (step - window.first);
        const std::vector<InjDesc> &all = window.all;
        const InjDesc *device_inj = window.device;
        if (accumulating) ensure_intensity();
        const int r = effective_rows();
        const bool pairs = kernel == DP_KERNEL_FUSED && (active().flags & DP_TILE_TWO_STEP);
        for (uint64_t s = 0; s < steps; ++s) {
            if (pairs && s + 1 < steps) {
                launch_pair(table, s, device_inj);
                check(hipGetLastError(), "kernel launch");
                ++launches;
                if (accumulating) accumulated += 2;
                ++s;
                continue;
            }
            const auto &t = table[s];
            if (kernel == DP_KERNEL_FUSED || kernel == DP_KERNEL_FUSED_OVERLAP) {
                StepArgs a = args(current, 1 - current);
                a.injections = device_inj + t.offset;
                a.n_magnetic = t.magnetic;
                a.n_electric = t.electric;
                a.intensity = accumulating ? intensity.as<double>() : nullptr;
                if (kernel == DP_KERNEL_FUSED) {
                    // a two-step tiling's odd last step takes its one-step shape
                    const Tiling tl = one_step_tiling(active());
                    const int rows = pairs ? tl.rows : r;
                    fill_tiles(a, rows, kAlignedX, tl);
                    launch_fused_aligned(a, rows, (tl.flags & DP_TILE_NONTEMPORAL) != 0,
                                         (tl.flags & DP_TILE_BUFFER) != 0, stream);
                } else {
                    fill_tiles(a, r, kOwnedX, active());
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
        // The injection kernels take their descriptors by value, so no
        // device table is needed here.
        auto table = build_injections(step, steps, -1, all);
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
