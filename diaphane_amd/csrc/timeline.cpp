// A keyframe ring in device memory: Timeline (src/timeline.rs:76-212) whose
// snapshots (E and H, 24 bytes per cell) never leave HBM. Recording and
// restoring are device-to-device copies on the sim's stream, so scrubbing
// back through a run costs no PCIe traffic and no host synchronization.
//
// Each slot carries two events: `ready` after the copy that filled it, and
// `used` after the last restore that read it. A restore waits on `ready` and
// an overwrite on `used`, so a ring shared by sims on different streams stays
// ordered.
#include <deque>

#include "engine.hpp"

using namespace dpa;

struct dp_timeline {
    uint64_t interval = 1;
    uint32_t capacity = 1;
    uint64_t cells = 0;  // grid size of the recorded frames (0: none yet)
    struct Slot {
        DeviceBuffer fields;  // E (3N floats) then H (3N floats)
        hipEvent_t ready = nullptr, used = nullptr;
        bool filled = false, read = false;
    };
    std::deque<Slot> slots;  // a deque: slots are never moved once made
    struct Frame {
        uint32_t slot;
        uint64_t step;
    };
    std::deque<Frame> ring;  // oldest first

    ~dp_timeline() { drop_slots(); }

    // Waits until no copy reads or fills a slot, then frees every slot and
    // its buffer.
    void drop_slots() noexcept {
        for (Slot &s : slots) {
            if (s.ready) (void)hipEventSynchronize(s.ready);
            if (s.used) (void)hipEventSynchronize(s.used);
            if (s.ready) (void)hipEventDestroy(s.ready);
            if (s.used) (void)hipEventDestroy(s.used);
            s.ready = s.used = nullptr;
        }
        slots.clear();
    }

    // Timeline::clear (src/timeline.rs:157-160) exists for when the scene
    // underneath has changed, so it also forgets the grid size the frames
    // were bound to and frees their buffers: the next keyframe may come from
    // a sim of another size.
    void clear() {
        ring.clear();
        drop_slots();
        cells = 0;
    }

    void bind(const dp_sim *sim) {
        if (cells == 0) {
            cells = sim->cells;
            return;
        }
        if (sim->cells != cells) fail(DP_ERR_SIZE, "snapshot was taken from a different grid");
    }

    uint32_t take_slot() {
        if (ring.size() == capacity) {
            const uint32_t slot = ring.front().slot;
            ring.pop_front();
            return slot;
        }
        std::vector<bool> busy(slots.size(), false);
        for (const Frame &f : ring) busy[f.slot] = true;
        for (uint32_t k = 0; k < slots.size(); ++k)
            if (!busy[k]) return k;
        slots.emplace_back();
        return static_cast<uint32_t>(slots.size() - 1);
    }

    void record(dp_sim *sim) {
        const uint32_t index = take_slot();
        Slot &s = slots[index];
        const size_t field_bytes = 3 * cells * sizeof(float);
        if (!s.fields.ptr) {
            s.fields.allocate(2 * field_bytes);
            check(hipEventCreateWithFlags(&s.ready, hipEventDisableTiming), "hipEventCreateWithFlags");
            check(hipEventCreateWithFlags(&s.used, hipEventDisableTiming), "hipEventCreateWithFlags");
        }
        if (s.read) check(hipStreamWaitEvent(sim->stream, s.used, 0), "hipStreamWaitEvent");
        char *base = static_cast<char *>(s.fields.ptr);
        check(hipMemcpyAsync(base, sim->e[sim->current].ptr, field_bytes, hipMemcpyDeviceToDevice, sim->stream),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(base + field_bytes, sim->h[sim->current].ptr, field_bytes, hipMemcpyDeviceToDevice,
                             sim->stream),
              "hipMemcpyAsync");
        check(hipEventRecord(s.ready, sim->stream), "hipEventRecord");
        s.filled = true;
        ring.push_back({index, sim->step});
    }

    void restore(dp_sim *sim, const Frame &frame) {
        Slot &s = slots[frame.slot];
        const size_t field_bytes = 3 * cells * sizeof(float);
        sim->clear_intensity();  // the average belongs to the abandoned history
        check(hipStreamWaitEvent(sim->stream, s.ready, 0), "hipStreamWaitEvent");
        const char *base = static_cast<const char *>(s.fields.ptr);
        check(hipMemcpyAsync(sim->e[sim->current].ptr, base, field_bytes, hipMemcpyDeviceToDevice, sim->stream),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(sim->h[sim->current].ptr, base + field_bytes, field_bytes, hipMemcpyDeviceToDevice,
                             sim->stream),
              "hipMemcpyAsync");
        check(hipEventRecord(s.used, sim->stream), "hipEventRecord");
        s.read = true;
        sim->step = frame.step;
    }
};

namespace {
void check_status(int status) {
    if (status != DP_OK) throw Error{status, last_error()};
}

void require_solo(const dp_sim *sim) {
    if (!sim) fail(DP_ERR_INVALID, "null sim");
    if (sim->batch) fail(DP_ERR_INVALID, "this sim is advanced by its batch");
}
}  // namespace

extern "C" {

int dp_timeline_create(uint64_t interval, uint32_t capacity, dp_timeline **out) {
    return guarded([&] {
        if (!out) fail(DP_ERR_INVALID, "null argument");
        if (interval == 0) fail(DP_ERR_INVALID, "keyframe interval must be positive");
        if (capacity == 0) fail(DP_ERR_INVALID, "a timeline needs room for at least one keyframe");
        auto t = std::make_unique<dp_timeline>();
        t->interval = interval;
        t->capacity = capacity;
        *out = t.release();
    });
}

void dp_timeline_destroy(dp_timeline *timeline) { delete timeline; }

int dp_timeline_observe(dp_timeline *t, dp_sim *sim) {
    return guarded([&] {
        if (!t) fail(DP_ERR_INVALID, "null timeline");
        require_solo(sim);
        const uint64_t step = sim->step;
        uint64_t newest = 0;
        if (!t->ring.empty()) {
            // Territory already covered (replaying after a backward seek):
            // recording here would break the ordering the backward scan needs.
            if (step <= t->ring.back().step) return;
            newest = t->ring.back().step / t->interval;
        }
        if (step / t->interval <= newest) return;
        t->bind(sim);
        t->record(sim);
    });
}

int dp_timeline_seek(dp_timeline *t, dp_sim *sim, uint64_t target, int *kind, uint64_t *replayed, uint64_t *at) {
    return guarded([&] {
        if (!t) fail(DP_ERR_INVALID, "null timeline");
        require_solo(sim);
        const uint64_t current = sim->step;
        int k = DP_SEEK_UNCHANGED;
        uint64_t r = 0, a = 0;
        if (target > current) {
            check_status(dp_sim_advance(sim, target - current));
            k = DP_SEEK_FORWARD;
            r = target - current;
        } else if (target < current) {
            const dp_timeline::Frame *found = nullptr;
            for (auto it = t->ring.rbegin(); it != t->ring.rend(); ++it)
                if (it->step <= target) {
                    found = &*it;
                    break;
                }
            if (found) {
                t->bind(sim);
                const uint64_t from = found->step;
                t->restore(sim, *found);
                check_status(dp_sim_advance(sim, target - from));
                k = DP_SEEK_FROM_KEYFRAME;
                r = target - from;
                a = from;
            } else {
                check_status(dp_sim_reset(sim));
                check_status(dp_sim_advance(sim, target));
                k = DP_SEEK_FROM_START;
                r = target;
            }
        }
        if (kind) *kind = k;
        if (replayed) *replayed = r;
        if (at) *at = a;
    });
}

int dp_timeline_clear(dp_timeline *t) {
    return guarded([&] {
        if (!t) fail(DP_ERR_INVALID, "null timeline");
        t->clear();
    });
}

int dp_timeline_info(const dp_timeline *t, uint32_t *keyframes, uint64_t *earliest, uint64_t *bytes) {
    return guarded([&] {
        if (!t) fail(DP_ERR_INVALID, "null timeline");
        if (keyframes) *keyframes = static_cast<uint32_t>(t->ring.size());
        if (earliest) *earliest = t->ring.empty() ? 0 : t->ring.front().step;
        if (bytes) *bytes = uint64_t(t->ring.size()) * 24 * t->cells;
    });
}

}  // extern "C"
