// The batched form: many independent sims of uneven sizes advanced together
// by one kernel launch per step, or per two steps, over a packed instance
// table.
//
// A batch borrows ordinary dp_sims. While a sim is a member, the batch
// advances it (its own advance entry point refuses) and its results are read
// through the ordinary dp_sim_read_* calls. Each member may carry a target
// step; a member at its target sits out the remaining launches, which is
// what lets a server admit and retire requests between chunks.
//
// A chunk of K steps is prepared on the host — per step and member a
// StepSlot (active, ping-pong parity, injections) and the injection list —
// uploaded in one copy, then launched as K kernel nodes of a captured HIP
// graph (re-captured only when the membership or K changes). Two staging sets
// alternate so the host prepares chunk n+1 while the device runs chunk n.
#include "engine.hpp"

#include <algorithm>
#include <limits>

namespace {

struct StagingSet {
    PinnedBuffer host;        // slots then injections
    DeviceBuffer slots;       // K × members StepSlot
    DeviceBuffer injections;  // InjDesc list of the chunk
    hipEvent_t done = nullptr;
    bool in_flight = false;
    hipGraphExec_t graph = nullptr;
    uint32_t graph_steps = 0;
    uint64_t graph_version = 0;

    void drop_graph() {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        graph_steps = 0;
    }

    // Makes room for a chunk of `slot_bytes` of slots and `inj_bytes` of
    // injections. The caller has waited for this set's previous chunk, so
    // no copy or launch reads these buffers. They only grow, with headroom,
    // and a graph captured from this set (it holds the device addresses as
    // kernel arguments) is destroyed before any buffer it names is freed.
    void reserve(size_t slot_bytes, size_t inj_bytes) {
        const bool grow_slots = slots.bytes < slot_bytes;
        const bool grow_injections = injections.bytes < inj_bytes;
        if (grow_slots || grow_injections) drop_graph();
        if (grow_slots) slots.allocate(slot_bytes * 2);
        if (grow_injections) injections.allocate(inj_bytes * 2);
        host.reserve(slot_bytes + inj_bytes);
    }
};

}  // namespace

struct dp_batch {
    dp_engine *engine = nullptr;
    hipStream_t stream = nullptr;
    std::vector<dp_sim *> members;
    std::vector<uint64_t> targets;
    bool dirty = true;
    uint64_t version = 0;

    // Defaults measured on MI355X over 256 uneven grids (48–160 cells a
    // side): two steps per launch with 8 rows, buffer addressing and cached
    // stores is the autotuner's pick (1.84–1.89 ms a step against 1.95–1.98
    // with 6 rows; on the served stream's 200 grids, 32–256 a side, 2.35
    // against 2.40 ms), 1.6x the best one-step form (one wave per tile, 4
    // rows, 16-plane slabs). Workgroups go to the XCDs in runs of 4 (16
    // consecutive tiles, y-neighbours sharing one L2): 1.5% faster than the
    // hardware's round robin there (runs of 2, 8 and 16: 0.9–1.2%).
    // depth 0 sizes the slabs to the membership (auto_depth): a batch of one
    // small grid needs short slabs to put enough waves on the device.
    int rows = 8;
    uint32_t depth = 0;
    uint32_t slab = 16;  // the slab depth the current tables use
    int flags = DP_BATCH_WAVE_PER_TILE | DP_TILE_TWO_STEP | DP_TILE_BUFFER | DP_TILE_XCD_REMAP | (4 << 8);
    uint32_t graph_steps = 32;

    DeviceBuffer instance_table;
    DeviceBuffer owners;
    std::vector<const void *> table_intensity;  // per member, as the table holds it
    DeviceBuffer trial_slots, trial_injections;
    uint32_t n_tiles = 0;
    uint32_t blocks = 0;
    StagingSet sets[2];
    int next_set = 0;
    hipEvent_t events[8] = {};
    uint64_t launches = 0;

    ~dp_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (dp_sim *sim : members) restore(sim);
        for (auto &s : sets) {
            s.drop_graph();
            if (s.done) (void)hipEventDestroy(s.done);
        }
        for (auto &e : events)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    static void restore(dp_sim *sim) {
        sim->stream = sim->own_stream;
        sim->own_stream = nullptr;
        sim->batch = nullptr;
    }

    // Removes a member. Never throws: it runs from dp_sim's destructor. A
    // failed wait here leaves the error to the stream's next checked call;
    // dp_batch_remove waits (checked) before it gets here.
    void detach(dp_sim *sim) noexcept {
        const size_t i = index_of(sim);
        if (i == members.size()) return;
        (void)hipStreamSynchronize(stream);
        for (auto &set : sets) set.in_flight = false;
        members.erase(members.begin() + static_cast<std::ptrdiff_t>(i));
        targets.erase(targets.begin() + static_cast<std::ptrdiff_t>(i));
        restore(sim);
        dirty = true;
    }

    // Two steps per launch use one wave per tile, or with DP_TILE_STACK one
    // workgroup of stacked waves per tile.
    bool two_step() const { return (flags & DP_TILE_TWO_STEP) != 0; }
    bool stack() const { return two_step() && (flags & DP_TILE_STACK) != 0; }
    int waves() const { return dp_sim::stack_waves(flags); }
    bool direct() const { return (flags & DP_BATCH_WAVE_PER_TILE) != 0 || two_step(); }
    // Workgroups per XCD run for the two-step kernel (DP_TILE_XCD_REMAP,
    // group in bits 8..15, 0 meaning 8); 0 without the flag.
    uint32_t xcd_group() const {
        if (!(flags & DP_TILE_XCD_REMAP)) return 0;
        const uint32_t group = (uint32_t(flags) >> 8) & 0xFFu;
        return group ? group : 8;
    }
    // Launches that advance a member by `steps` steps.
    uint32_t launches_for(uint32_t steps) const { return two_step() ? (steps + 1) / 2 : steps; }
    // Buffer addressing when asked for and every member's offsets fit it.
    bool buffer() const {
        if (!(flags & DP_TILE_BUFFER)) return false;
        for (const dp_sim *s : members)
            if (!buffer_addressing_fits(s->cells, s->grid.extent[0], s->grid.extent[1])) return false;
        return true;
    }

    size_t index_of(const dp_sim *sim) const {
        for (size_t i = 0; i < members.size(); ++i)
            if (members[i] == sim) return i;
        return members.size();
    }

    void sync() { check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

    uint32_t owned_x() const { return two_step() ? kPairOwnedX : kAlignedX; }
    int owned_rows() const { return stack() ? waves() * rows - 4 : rows - (two_step() ? 4 : 2); }
    // Column and row tiles of a member's grid (stacked two-step tiles own
    // more rows at the walls).
    uint32_t tiles_x_of(uint32_t nx) const { return (nx + owned_x() - 1) / owned_x(); }
    uint32_t tiles_y_of(uint32_t ny) const {
        return stack() ? stack_tiles_y(ny, uint32_t(waves() * rows)) : (ny + owned_rows() - 1) / owned_rows();
    }

    void tiles_of(const dp_sim *sim, InstanceDesc &d) const {
        const uint32_t nx = sim->grid.extent[0], ny = sim->grid.extent[1], nz = sim->grid.extent[2];
        d.tiles_x = tiles_x_of(nx);
        d.tiles_y = tiles_y_of(ny);
        d.depth = std::min<uint32_t>(nz, slab);
        d.tiles_z = (nz + d.depth - 1) / d.depth;
        d.tiles_total = d.tiles_x * d.tiles_y * d.tiles_z;
    }

    // The slab depth for depth 0: the single-sim rule applied to the x-y
    // tiles of every member together, at most 16 for one step per launch and
    // 32 for two (the measured best for a full batch).
    uint32_t auto_depth() const {
        uint64_t tiles_xy = 0;
        for (const dp_sim *s : members)
            tiles_xy += uint64_t(tiles_x_of(s->grid.extent[0])) * tiles_y_of(s->grid.extent[1]);
        if (stack()) tiles_xy *= uint64_t(waves());  // the rule counts waves
        const uint32_t capped = static_cast<uint32_t>(std::min<uint64_t>(tiles_xy, 0xFFFFFFFFull));
        return two_step() ? std::min<uint32_t>(32, dp_sim::default_pair_depth(capped))
                          : std::min<uint32_t>(16, dp_sim::default_depth(capped));
    }

    // Instance table and tile owners for the current membership.
    void rebuild() {
        sync();
        slab = depth ? depth : auto_depth();
        std::vector<InstanceDesc> table(members.size());
        std::vector<uint32_t> owner;
        table_intensity.assign(members.size(), nullptr);
        for (dp_sim *sim : members)
            if (sim->accumulating) sim->ensure_intensity();
        for (size_t i = 0; i < members.size(); ++i) {
            dp_sim *s = members[i];
            InstanceDesc &d = table[i];
            d = InstanceDesc{};
            d.e[0] = s->e[0].as<float>();
            d.e[1] = s->e[1].as<float>();
            d.h[0] = s->h[0].as<float>();
            d.h[1] = s->h[1].as<float>();
            d.material = s->material.as<uint8_t>();
            d.coefficients = s->coefficients.as<float4>();
            d.axis_table = s->axis_table.as<float>();
            d.weights = s->weights.as<float>();
            d.intensity = s->intensity.as<double>();
            d.nx = s->grid.extent[0];
            d.ny = s->grid.extent[1];
            d.nz = s->grid.extent[2];
            d.material_count = s->material_count;
            d.cells = s->cells;
            tiles_of(s, d);
            d.tile_base = static_cast<uint32_t>(owner.size());
            table_intensity[i] = s->intensity.ptr;
            owner.insert(owner.end(), d.tiles_total, static_cast<uint32_t>(i));
        }
        if (owner.size() > std::numeric_limits<uint32_t>::max() / 2)
            fail(DP_ERR_UNSUPPORTED, "too many tiles in one batch");
        n_tiles = static_cast<uint32_t>(owner.size());
        // Both sets' graphs name the table and the owner list: they go before
        // either buffer can be freed (the tables only grow, with headroom).
        const size_t table_bytes = std::max<size_t>(table.size(), 1) * sizeof(InstanceDesc);
        const size_t owner_bytes = std::max<size_t>(owner.size(), 1) * sizeof(uint32_t);
        if (instance_table.bytes < table_bytes || owners.bytes < owner_bytes)
            for (auto &set : sets) set.drop_graph();
        instance_table.reserve(table_bytes);
        owners.reserve(owner_bytes);
        if (!table.empty())
            copy_on(stream, instance_table.ptr, table.data(), table.size() * sizeof(InstanceDesc), hipMemcpyHostToDevice);
        if (!owner.empty())
            copy_on(stream, owners.ptr, owner.data(), owner.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        const uint32_t resident = batch_resident_blocks(rows, (flags & DP_TILE_NONTEMPORAL) != 0);
        blocks = std::max<uint32_t>(1, std::min<uint32_t>(resident, (n_tiles + 3) / 4));
        ++version;
        dirty = false;
    }

    // Times `trials` launches of the current tiling without changing any
    // member: every member reads its current buffers and writes the other
    // half of its ping-pong pair (which the next real step overwrites), with
    // no injections and no intensity.
    float time_tiling(uint32_t trials) {
        if (dirty) rebuild();
        std::vector<StepSlot> slot(members.size());
        for (size_t i = 0; i < members.size(); ++i) {
            slot[i] = StepSlot{};
            slot[i].active = 1;
            slot[i].parity = static_cast<uint32_t>(members[i]->current);
        }
        trial_slots.reserve(std::max<size_t>(slot.size(), 1) * sizeof(StepSlot));
        copy_on(stream, trial_slots.ptr, slot.data(), slot.size() * sizeof(StepSlot), hipMemcpyHostToDevice);
        if (!trial_injections.ptr) trial_injections.allocate(sizeof(InjDesc));
        const bool nt = (flags & DP_TILE_NONTEMPORAL) != 0;
        auto launch = [&] {
            if (stack())
                launch_fused_stack_batch(instance_table.as<InstanceDesc>(), trial_slots.as<StepSlot>(),
                                         owners.as<uint32_t>(), n_tiles, trial_injections.as<InjDesc>(), rows, waves(),
                                         buffer(), xcd_group(), stream);
            else if (two_step())
                launch_fused_pair_batch(instance_table.as<InstanceDesc>(), trial_slots.as<StepSlot>(),
                                        owners.as<uint32_t>(), n_tiles, trial_injections.as<InjDesc>(), rows, nt,
                                        buffer(), xcd_group(), stream);
            else
                launch_fused_batch(instance_table.as<InstanceDesc>(), trial_slots.as<StepSlot>(), owners.as<uint32_t>(),
                                   n_tiles, trial_injections.as<InjDesc>(), rows, nt, buffer(), direct(), blocks,
                                   stream);
        };
        launch();  // first launch of a shape loads its code object
        check(hipGetLastError(), "kernel launch");
        check(hipEventRecord(events[6], stream), "hipEventRecord");
        for (uint32_t t = 0; t < trials; ++t) launch();
        check(hipEventRecord(events[7], stream), "hipEventRecord");
        check(hipEventSynchronize(events[7]), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, events[6], events[7]), "hipEventElapsedTime");
        return ms / static_cast<float>(two_step() ? 2 * trials : trials);  // per step
    }

    // Picks the grid form, rows and slab depth by timing every candidate on
    // the members as they are. Results never depend on the choice.
    void autotune(uint32_t trials, int *rows_out, int *depth_out, float *ms_out) {
        static const int row_choices[] = {4, 6, 8};
        static const int pair_row_choices[] = {6, 8, 10};
        static const int stack_row_choices[] = {4, 6, 6};  // (6 twice: the sweep takes three)
        static const uint32_t depth_choices[] = {8, 16, 32, 64};
        // One step per launch keeps non-temporal stores (measured faster);
        // two steps per launch try both store kinds.
        static const int grid_choices[] = {
            DP_TILE_NONTEMPORAL,
            DP_TILE_NONTEMPORAL | DP_BATCH_WAVE_PER_TILE,
            DP_TILE_NONTEMPORAL | DP_TILE_BUFFER,
            DP_TILE_NONTEMPORAL | DP_BATCH_WAVE_PER_TILE | DP_TILE_BUFFER,
            DP_TILE_TWO_STEP,
            DP_TILE_TWO_STEP | DP_TILE_BUFFER,
            DP_TILE_TWO_STEP | DP_TILE_NONTEMPORAL,
            DP_TILE_TWO_STEP | DP_TILE_NONTEMPORAL | DP_TILE_BUFFER,
            DP_TILE_TWO_STEP | DP_TILE_STACK | DP_TILE_BUFFER | (2 << 16),
            DP_TILE_TWO_STEP | DP_TILE_STACK | DP_TILE_BUFFER | (4 << 16)};
        const int tuned_bits =
            DP_BATCH_WAVE_PER_TILE | DP_TILE_BUFFER | DP_TILE_TWO_STEP | DP_TILE_NONTEMPORAL | DP_TILE_STACK | 0xF0000;
        int best_rows = rows, best_grid = flags & tuned_bits;
        uint32_t best_depth = depth;
        float best = std::numeric_limits<float>::infinity();
        struct Candidate {
            int grid, rows;
            uint32_t depth;
            float ms;
        };
        auto time_candidate = [&](const Candidate &c, uint32_t launches) {
            flags = (flags & ~tuned_bits) | c.grid;
            rows = c.rows;
            depth = c.depth;
            dirty = true;
            return time_tiling(launches);
        };
        if (!members.empty()) {
            std::vector<Candidate> sweep;
            for (int g : grid_choices)
                for (int k = 0; k < 3; ++k)
                    for (uint32_t d : depth_choices) {
                        if ((g & DP_TILE_STACK) && k == 2) continue;  // stacked tiles: rows 4 and 6
                        const int r = (g & DP_TILE_STACK)      ? stack_row_choices[k]
                                      : (g & DP_TILE_TWO_STEP) ? pair_row_choices[k]
                                                               : row_choices[k];
                        Candidate c{g, r, d, 0.0f};
                        c.ms = time_candidate(c, trials);
                        sweep.push_back(c);
                    }
            // The sweep's four fastest, timed again with three times the
            // launches in two alternating rounds: neighbouring shapes differ
            // by less than one short trial's noise.
            std::sort(sweep.begin(), sweep.end(), [](const Candidate &x, const Candidate &y) { return x.ms < y.ms; });
            sweep.resize(std::min<size_t>(sweep.size(), 4));
            for (Candidate &c : sweep) c.ms = std::numeric_limits<float>::infinity();
            for (int round = 0; round < 2; ++round)
                for (Candidate &c : sweep) c.ms = std::min(c.ms, time_candidate(c, 3 * trials));
            for (const Candidate &c : sweep)
                if (c.ms < best) {
                    best = c.ms;
                    best_rows = c.rows;
                    best_depth = c.depth;
                    best_grid = c.grid;
                }
            // The winner's slab depth between the powers of two, as for a
            // single sim.
            const uint32_t winner_depth = best_depth;
            for (uint32_t d : {winner_depth * 3 / 4, winner_depth * 3 / 2}) {
                if (d < 2 || d == winner_depth) continue;
                Candidate c{best_grid, best_rows, d, std::numeric_limits<float>::infinity()};
                for (int round = 0; round < 2; ++round) c.ms = std::min(c.ms, time_candidate(c, 3 * trials));
                if (c.ms < best) {
                    best = c.ms;
                    best_depth = d;
                }
            }
        }
        flags = (flags & ~tuned_bits) | best_grid;
        rows = best_rows;
        depth = best_depth;
        dirty = true;
        for (auto &set : sets) set.drop_graph();
        if (rows_out) *rows_out = rows;
        if (depth_out) *depth_out = static_cast<int>(depth);
        if (ms_out) *ms_out = members.empty() ? 0.0f : best;
    }

    // `steps` launches (one or two steps each, by the tiling); a full chunk
    // (`capture`) runs as a captured graph.
    void launch_chunk(StagingSet &set, uint32_t steps, bool capture) {
        const StepSlot *slots = set.slots.as<StepSlot>();
        const InjDesc *inj = set.injections.as<InjDesc>();
        const bool nt = (flags & DP_TILE_NONTEMPORAL) != 0;
        auto launch_all = [&] {
            for (uint32_t s = 0; s < steps; ++s) {
                if (stack())
                    launch_fused_stack_batch(instance_table.as<InstanceDesc>(), slots + size_t(s) * members.size(),
                                             owners.as<uint32_t>(), n_tiles, inj, rows, waves(), buffer(),
                                             xcd_group(), stream);
                else if (two_step())
                    launch_fused_pair_batch(instance_table.as<InstanceDesc>(), slots + size_t(s) * members.size(),
                                            owners.as<uint32_t>(), n_tiles, inj, rows, nt, buffer(), xcd_group(),
                                            stream);
                else
                    launch_fused_batch(instance_table.as<InstanceDesc>(), slots + size_t(s) * members.size(),
                                       owners.as<uint32_t>(), n_tiles, inj, rows, nt, buffer(), direct(), blocks,
                                       stream);
            }
        };
        if (!capture) {
            launch_all();
            check(hipGetLastError(), "kernel launch");
            return;
        }
        if (!set.graph || set.graph_version != version || set.graph_steps != steps) {
            set.drop_graph();
            hipGraph_t graph = nullptr;
            // Relaxed: calls other threads make meanwhile (a server's builders
            // allocating, a client copying a result out) cannot invalidate it.
            check(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
            launch_all();
            check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
            const hipError_t made = hipGraphInstantiate(&set.graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            check(made, "hipGraphInstantiate");
            set.graph_steps = steps;
            set.graph_version = version;
        }
        check(hipGraphLaunch(set.graph, stream), "hipGraphLaunch");
    }

    // Advances every member by min(steps, target − step).
    void advance(uint64_t steps) {
        if (members.empty() || steps == 0) return;
        // A member that switched intensity on since the table was built
        // brings a new buffer: the table must carry it.
        for (size_t i = 0; i < members.size() && !dirty; ++i) {
            if (members[i]->accumulating) members[i]->ensure_intensity();
            dirty = members[i]->intensity.ptr != table_intensity[i];
        }
        if (dirty) rebuild();
        const size_t n = members.size();
        std::vector<StepSlot> slot_buf;
        std::vector<InjDesc> inj_all, inj_one;
        while (steps > 0) {
            bool any = false;
            for (size_t i = 0; i < n; ++i) any = any || members[i]->step < targets[i];
            if (!any) return;
            const uint32_t chunk = static_cast<uint32_t>(
                std::min<uint64_t>(steps, graph_steps ? graph_steps : 64));
            const uint32_t chunk_launches = launches_for(chunk);
            slot_buf.assign(size_t(chunk_launches) * n, StepSlot{});
            inj_all.clear();
            for (size_t i = 0; i < n; ++i) {
                dp_sim *sim = members[i];
                const uint64_t room = targets[i] > sim->step ? targets[i] - sim->step : 0;
                const uint32_t active = static_cast<uint32_t>(std::min<uint64_t>(room, chunk));
                if (active == 0) continue;
                auto table = sim->build_injections(sim->step, active, +1, inj_one);
                const uint32_t base = static_cast<uint32_t>(inj_all.size());
                int parity = sim->current;
                // Launch s advances the member by steps s (one step per
                // launch), or by steps 2s and 2s + 1 (two), the last of an
                // odd count alone.
                const uint32_t per_launch = two_step() ? 2 : 1;
                for (uint32_t s = 0; s < launches_for(active); ++s) {
                    const uint32_t first = s * per_launch;
                    StepSlot &slot = slot_buf[size_t(s) * n + i];
                    slot.active = 1;
                    slot.parity = static_cast<uint32_t>(parity);
                    slot.inj_offset = base + table[first].offset;
                    slot.n_magnetic = table[first].magnetic;
                    slot.n_electric = table[first].electric;
                    slot.accumulate = sim->accumulating ? 1u : 0u;
                    if (two_step()) {
                        if (first + 1 < active) {
                            slot.inj_offset2 = base + table[first + 1].offset;
                            slot.n_magnetic2 = table[first + 1].magnetic;
                            slot.n_electric2 = table[first + 1].electric;
                        } else {
                            slot.single = 1;
                        }
                    }
                    parity = 1 - parity;
                }
                inj_all.insert(inj_all.end(), inj_one.begin(), inj_one.end());
            }
            StagingSet &set = sets[next_set];
            next_set ^= 1;
            if (set.in_flight) {
                check(hipEventSynchronize(set.done), "hipEventSynchronize");
                set.in_flight = false;
            }
            const size_t slot_bytes = slot_buf.size() * sizeof(StepSlot);
            const size_t inj_bytes = std::max<size_t>(inj_all.size(), 1) * sizeof(InjDesc);
            set.reserve(slot_bytes, inj_bytes);
            std::memcpy(set.host.ptr, slot_buf.data(), slot_bytes);
            if (!inj_all.empty())
                std::memcpy(static_cast<char *>(set.host.ptr) + slot_bytes, inj_all.data(),
                            inj_all.size() * sizeof(InjDesc));
            check(hipMemcpyAsync(set.slots.ptr, set.host.ptr, slot_bytes, hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync");
            check(hipMemcpyAsync(set.injections.ptr, static_cast<char *>(set.host.ptr) + slot_bytes, inj_bytes,
                                 hipMemcpyHostToDevice, stream),
                  "hipMemcpyAsync");
            launch_chunk(set, chunk_launches, graph_steps != 0 && chunk == graph_steps);
            check(hipEventRecord(set.done, stream), "hipEventRecord");
            set.in_flight = true;
            launches += chunk_launches;
            for (size_t i = 0; i < n; ++i) {
                dp_sim *sim = members[i];
                const uint64_t room = targets[i] > sim->step ? targets[i] - sim->step : 0;
                const uint32_t active = static_cast<uint32_t>(std::min<uint64_t>(room, chunk));
                const uint32_t member_launches = launches_for(active);
                sim->step += active;
                if (member_launches & 1) sim->current = 1 - sim->current;  // each launch flips the pair once
                if (sim->accumulating) sim->accumulated += active;
                sim->launches += member_launches;
            }
            steps -= chunk;
        }
    }

};

extern "C" {

int dp_batch_create(dp_engine *engine, dp_batch **out) {
    return guarded([&] {
        if (!engine || !out) fail(DP_ERR_INVALID, "null argument");
        check(hipSetDevice(engine->device), "hipSetDevice");
        auto batch = std::make_unique<dp_batch>();
        batch->engine = engine;
        check(hipStreamCreateWithFlags(&batch->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto &s : batch->sets) check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate");
        for (auto &e : batch->events) check(hipEventCreate(&e), "hipEventCreate");
        *out = batch.release();
    });
}

void dp_batch_destroy(dp_batch *batch) { delete batch; }

int dp_batch_add(dp_batch *batch, dp_sim *sim) {
    return guarded([&] {
        if (!batch || !sim) fail(DP_ERR_INVALID, "null argument");
        if (sim->batch) fail(DP_ERR_INVALID, "the sim already belongs to a batch");
        if (sim->cells * sizeof(float) > 0xFFFFFFFFull)
            fail(DP_ERR_UNSUPPORTED, "a batched grid must fit 4 GiB per field component");
        sim->sync();
        batch->sync();
        sim->batch = batch;
        sim->own_stream = sim->stream;
        sim->stream = batch->stream;
        batch->members.push_back(sim);
        batch->targets.push_back(std::numeric_limits<uint64_t>::max());
        batch->dirty = true;
    });
}

int dp_batch_remove(dp_batch *batch, dp_sim *sim) {
    return guarded([&] {
        if (!batch || !sim) fail(DP_ERR_INVALID, "null argument");
        const size_t i = batch->index_of(sim);
        if (i == batch->members.size()) fail(DP_ERR_INVALID, "the sim is not a member of this batch");
        batch->sync();  // a pending device error surfaces here, as a status
        batch->detach(sim);
    });
}

int dp_batch_set_target(dp_batch *batch, dp_sim *sim, uint64_t step) {
    return guarded([&] {
        if (!batch || !sim) fail(DP_ERR_INVALID, "null argument");
        const size_t i = batch->index_of(sim);
        if (i == batch->members.size()) fail(DP_ERR_INVALID, "the sim is not a member of this batch");
        batch->targets[i] = step;
    });
}

int dp_batch_advance(dp_batch *batch, uint64_t steps) {
    return guarded([&] {
        if (!batch) fail(DP_ERR_INVALID, "null batch");
        batch->advance(steps);
    });
}

int dp_batch_wait(dp_batch *batch, uint32_t timeout_ms) {
    return guarded([&] {
        if (!batch) fail(DP_ERR_INVALID, "null batch");
        const uint32_t limit = timeout_ms ? timeout_ms : 120000;
        const auto start = std::chrono::steady_clock::now();
        for (;;) {
            hipError_t status = hipStreamQuery(batch->stream);
            if (status == hipSuccess) return;
            if (status != hipErrorNotReady) check(status, "hipStreamQuery");
            if (std::chrono::steady_clock::now() - start > std::chrono::milliseconds(limit))
                fail(DP_ERR_TIMEOUT, "the GPU did not finish in time");
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    });
}

int dp_batch_set_graph_steps(dp_batch *batch, uint32_t steps) {
    return guarded([&] {
        if (!batch) fail(DP_ERR_INVALID, "null batch");
        if (steps > 4096) fail(DP_ERR_INVALID, "at most 4096 steps per graph");
        batch->graph_steps = steps;
    });
}

int dp_batch_autotune(dp_batch *batch, uint32_t trials, int *rows, int *depth, float *ms_per_step) {
    return guarded([&] {
        if (!batch) fail(DP_ERR_INVALID, "null batch");
        if (trials == 0 || trials > 1000) fail(DP_ERR_INVALID, "trials must be in 1..1000");
        batch->sync();
        for (auto &set : batch->sets) set.in_flight = false;
        batch->autotune(trials, rows, depth, ms_per_step);
    });
}

int dp_batch_set_tile(dp_batch *batch, int rows, int depth, int flags) {
    return guarded([&] {
        if (!batch) fail(DP_ERR_INVALID, "null batch");
        if (depth < 0) fail(DP_ERR_INVALID, "depth must be non-negative");
        const int new_flags =
            flags >= 0 ? flags & (DP_TILE_NONTEMPORAL | DP_TILE_BUFFER | DP_BATCH_WAVE_PER_TILE | DP_TILE_TWO_STEP |
                                  DP_TILE_XCD_REMAP | DP_TILE_STACK | 0xFF00 | 0xF0000)
                       : batch->flags;
        const bool two = (new_flags & DP_TILE_TWO_STEP) != 0;
        const bool stacked = two && (new_flags & DP_TILE_STACK) != 0;
        if (rows != 0 && (stacked ? rows != 4 && rows != 6
                          : two   ? rows != 6 && rows != 8 && rows != 10
                                  : rows != 4 && rows != 6 && rows != 8))
            fail(DP_ERR_INVALID, stacked ? "rows must be 4 or 6 for stacked waves"
                                 : two   ? "rows must be 6, 8 or 10 for two steps per launch"
                                         : "rows must be 4, 6 or 8");
        batch->sync();
        batch->flags = new_flags;
        batch->rows = rows ? rows : (stacked ? 4 : two ? 8 : 4);
        batch->depth = static_cast<uint32_t>(depth);
        batch->dirty = true;
    });
}

int dp_batch_tiling(const dp_batch *batch, int *rows, int *depth, int *flags) {
    return guarded([&] {
        if (!batch) fail(DP_ERR_INVALID, "null batch");
        if (rows) *rows = batch->rows;
        if (depth) *depth = static_cast<int>(batch->depth ? batch->depth : batch->auto_depth());
        if (flags) *flags = batch->flags;
    });
}

uint32_t dp_batch_size(const dp_batch *batch) { return batch ? static_cast<uint32_t>(batch->members.size()) : 0; }

uint64_t dp_batch_cells(const dp_batch *batch) {
    uint64_t total = 0;
    if (batch)
        for (const dp_sim *s : batch->members) total += s->cells;
    return total;
}

uint64_t dp_batch_launch_count(const dp_batch *batch) { return batch ? batch->launches : 0; }

int dp_batch_record_event(dp_batch *batch, int slot) {
    return guarded([&] {
        if (!batch || slot < 0 || slot >= 8) fail(DP_ERR_INVALID, "bad event slot");
        check(hipEventRecord(batch->events[slot], batch->stream), "hipEventRecord");
    });
}

int dp_batch_elapsed_ms(dp_batch *batch, int from_slot, int to_slot, float *ms) {
    return guarded([&] {
        if (!batch || !ms || from_slot < 0 || from_slot >= 8 || to_slot < 0 || to_slot >= 8)
            fail(DP_ERR_INVALID, "bad event slot");
        check(hipEventSynchronize(batch->events[to_slot]), "hipEventSynchronize");
        check(hipEventElapsedTime(ms, batch->events[from_slot], batch->events[to_slot]), "hipEventElapsedTime");
    });
}

}  // extern "C"

dp_sim::~dp_sim() {
    if (batch) batch->detach(this);
    release_stream();
}
