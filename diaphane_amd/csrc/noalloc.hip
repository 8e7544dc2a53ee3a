// The viewer's auto-range reduction without any allocation
// (dp_sim_measure_peak_noalloc), for callers that cannot allocate: code
// under HIP graph capture, tight serving loops, harnesses that forbid it.
//
// dp_sim_measure_peak allocates its per-block partials on a sim's first call
// and reduces them on the host. This variant keeps the partials and the final
// eight maxima in a file-scope device array of fixed size, one per GPU for
// the life of the process, reduces on the device and copies 32 bytes to the
// caller. Calls that overlap on one GPU share that array and race; every
// call writes all of the array it reads, so nothing needs clearing.
#include "engine.hpp"

#include <mutex>

namespace dpa {
namespace {

// The largest partial count the peak kernel uses (peak_blocks clamps to it).
constexpr uint32_t kScratchBlocks = 4096;
constexpr uint32_t kScratchSlots = 8;
constexpr uint32_t kScratchFloats = kScratchBlocks * kScratchSlots + kScratchSlots;  // partials, then maxima

static __device__ __align__(256) float scratch[kScratchFloats];

// The device address of `scratch` on each device, taken once per device at
// engine creation (the array's name is not a device address in host code).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
float *g_scratch[kMaxDevices] = {};

// The per-block maxima of k_measure_peak into eight: the same fmaxf over
// the same values as the host loop of dp_sim_measure_peak (max is exact and
// independent of order), from 0.
__global__ __launch_bounds__(256) void k_peak_final(const float *partials, uint32_t blocks, float *out) {
    __shared__ float s_max[kScratchSlots][256];
    float m[kScratchSlots] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
#pragma unroll
        for (uint32_t k = 0; k < kScratchSlots; ++k) m[k] = fmaxf(m[k], partials[size_t(b) * kScratchSlots + k]);
#pragma unroll
    for (uint32_t k = 0; k < kScratchSlots; ++k) s_max[k][threadIdx.x] = m[k];
    __syncthreads();
    for (uint32_t stride = blockDim.x / 2; stride > 0; stride /= 2) {
        if (threadIdx.x < stride)
#pragma unroll
            for (uint32_t k = 0; k < kScratchSlots; ++k)
                s_max[k][threadIdx.x] = fmaxf(s_max[k][threadIdx.x], s_max[k][threadIdx.x + stride]);
        __syncthreads();
    }
    if (threadIdx.x < kScratchSlots) out[threadIdx.x] = s_max[threadIdx.x][0];
}

}  // namespace

// Called by dp_engine_create with `device` current.
void prepare_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error{DP_ERR_UNSUPPORTED, "device index beyond 63"};
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (g_scratch[device]) return;
    void *ptr = nullptr;
    const hipError_t status = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (status != hipSuccess || !ptr) throw HipFailure{status, "hipGetSymbolAddress"};
    g_scratch[device] = static_cast<float *>(ptr);
}

}  // namespace dpa

extern "C" int dp_sim_measure_peak_noalloc(dp_sim *sim, float peak[8]) {
    return guarded([&] {
        if (!sim || !peak) fail(DP_ERR_INVALID, "null argument");
        const int device = sim->engine->device;
        float *base = nullptr;
        {
            std::lock_guard<std::mutex> lock(dpa::g_scratch_lock);
            if (device >= 0 && device < dpa::kMaxDevices) base = dpa::g_scratch[device];
        }
        if (!base) fail(DP_ERR_DEVICE, "the device's scratch array was not prepared");
        const uint32_t blocks = peak_blocks(sim->cells);
        if (blocks > dpa::kScratchBlocks)
            fail(DP_ERR_SIZE, "the grid needs more peak partials than the fixed scratch array holds");
        const double *intensity = sim->intensity.ptr && sim->accumulated ? sim->intensity.as<double>() : nullptr;
        float *maxima = base + size_t(dpa::kScratchBlocks) * dpa::kScratchSlots;
        launch_measure_peak(sim->e[sim->current].as<float>(), sim->h[sim->current].as<float>(), intensity,
                            sim->accumulated, sim->cells, base, sim->stream);
        dpa::k_peak_final<<<1, 256, 0, sim->stream>>>(base, blocks, maxima);
        check(hipGetLastError(), "kernel launch");
        copy_on(sim->stream, peak, maxima, 8 * sizeof(float), hipMemcpyDeviceToHost);
    });
}
