// Two-step fused kernel with the waves of a workgroup stacked along y
// (DP_TILE_STACK), in its own translation unit so it compiles in parallel
// with the others. See fused_stack.inc for the tile and kernels.hpp for the
// layout and the arithmetic contract.
#include "kernels.hpp"

namespace dpa {

namespace {

#include "device_common.inc"
#include "fused_access.inc"
#include "fused_stack.inc"

}  // namespace

template <int W, template <bool> class Access, bool kIntensity>
static void launch_stack_as(const StepArgs &args, int rows, hipStream_t stream) {
    const uint32_t blocks = args.tiles_total;
    switch (rows) {
    case 4: k_fused_stack<4, W, Access<false>, kIntensity><<<blocks, W * kLanes, 0, stream>>>(args); break;
    default: k_fused_stack<6, W, Access<false>, kIntensity><<<blocks, W * kLanes, 0, stream>>>(args); break;
    }
}

template <template <bool> class Access>
static void launch_stack_with(const StepArgs &args, int rows, int waves, hipStream_t stream) {
    // Launches without an intensity sum use a build without its code.
    if (waves == 2) {
        if (args.intensity)
            launch_stack_as<2, Access, true>(args, rows, stream);
        else
            launch_stack_as<2, Access, false>(args, rows, stream);
    } else {
        if (args.intensity)
            launch_stack_as<4, Access, true>(args, rows, stream);
        else
            launch_stack_as<4, Access, false>(args, rows, stream);
    }
}

void launch_fused_stack(const StepArgs &args, int rows, int waves, bool buffer, hipStream_t stream) {
    if (args.tiles_total == 0) return;
    if (buffer && buffer_addressing_fits(args.cells, args.nx, args.ny))
        launch_stack_with<BufferAccess>(args, rows, waves, stream);
    else
        launch_stack_with<Pointer64>(args, rows, waves, stream);
}

template <int W, template <bool> class Access>
static void launch_stack_batch_as(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                                  uint32_t n_tiles, const InjDesc *injections, int rows, uint32_t xcd_group,
                                  hipStream_t stream) {
    switch (rows) {
    case 4:
        k_fused_stack_batch<4, W, Access<false>><<<n_tiles, W * kLanes, 0, stream>>>(instances, slots, owners,
                                                                                    n_tiles, injections, xcd_group);
        break;
    default:
        k_fused_stack_batch<6, W, Access<false>><<<n_tiles, W * kLanes, 0, stream>>>(instances, slots, owners,
                                                                                    n_tiles, injections, xcd_group);
        break;
    }
}

void launch_fused_stack_batch(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                              uint32_t n_tiles, const InjDesc *injections, int rows, int waves, bool buffer,
                              uint32_t xcd_group, hipStream_t stream) {
    if (n_tiles == 0) return;
    if (waves == 2) {
        if (buffer)
            launch_stack_batch_as<2, BufferAccess>(instances, slots, owners, n_tiles, injections, rows, xcd_group,
                                                   stream);
        else
            launch_stack_batch_as<2, Pointer64>(instances, slots, owners, n_tiles, injections, rows, xcd_group,
                                                stream);
    } else {
        if (buffer)
            launch_stack_batch_as<4, BufferAccess>(instances, slots, owners, n_tiles, injections, rows, xcd_group,
                                                   stream);
        else
            launch_stack_batch_as<4, Pointer64>(instances, slots, owners, n_tiles, injections, rows, xcd_group,
                                                stream);
    }
}

}  // namespace dpa
