// Two-step fused kernel for gfx950 (its own translation unit, so the
// single-step kernels and this one compile in parallel). See kernels.hpp for
// the layout and the arithmetic contract.
#include "kernels.hpp"

namespace dpa {

namespace {

#include "device_common.inc"
#include "fused_access.inc"
#include "fused_pair.inc"

}  // namespace

template <template <bool> class Access, bool kIntensity>
static void launch_fused_pair_as(const StepArgs &args, int rows, bool nontemporal, hipStream_t stream) {
    const uint32_t blocks = (args.tiles_total + 3) / 4;
#define DPA_ROWS(R)                                                                           \
    case R:                                                                                   \
        if (nontemporal)                                                                      \
            k_fused_pair<R, Access<true>, kIntensity><<<blocks, 256, 0, stream>>>(args);      \
        else                                                                                  \
            k_fused_pair<R, Access<false>, kIntensity><<<blocks, 256, 0, stream>>>(args);     \
        break;
    switch (rows) {
        DPA_ROWS(6)
        DPA_ROWS(8)
        DPA_ROWS(10)
    default:
        k_fused_pair<8, Access<true>, kIntensity><<<blocks, 256, 0, stream>>>(args);
        break;
    }
#undef DPA_ROWS
}

// Launches without an intensity sum use a build of the kernel without its
// read-modify-write code.
template <template <bool> class Access>
static void launch_fused_pair_with(const StepArgs &args, int rows, bool nontemporal, hipStream_t stream) {
    if (args.intensity)
        launch_fused_pair_as<Access, true>(args, rows, nontemporal, stream);
    else
        launch_fused_pair_as<Access, false>(args, rows, nontemporal, stream);
}

void launch_fused_pair(const StepArgs &args, int rows, bool nontemporal, bool buffer, hipStream_t stream) {
    if (buffer && buffer_addressing_fits(args.cells, args.nx, args.ny))
        launch_fused_pair_with<BufferAccess>(args, rows, nontemporal, stream);
    else if (args.cells * sizeof(float) <= 0xFFFFFFFFull)
        launch_fused_pair_with<Pointer32>(args, rows, nontemporal, stream);
    else
        launch_fused_pair_with<Pointer64>(args, rows, nontemporal, stream);
}

template <template <bool> class Access>
static void launch_pair_batch_as(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                                 uint32_t n_tiles, const InjDesc *injections, int rows, bool nontemporal,
                                 uint32_t xcd_group, hipStream_t stream) {
    const uint32_t blocks = (n_tiles + 3) / 4;
#define DPA_ROWS(R)                                                                                             \
    case R:                                                                                                     \
        if (nontemporal)                                                                                        \
            k_fused_pair_batch<R, Access<true>><<<blocks, 256, 0, stream>>>(instances, slots, owners, n_tiles,   \
                                                                           injections, xcd_group);              \
        else                                                                                                    \
            k_fused_pair_batch<R, Access<false>><<<blocks, 256, 0, stream>>>(instances, slots, owners, n_tiles,  \
                                                                            injections, xcd_group);             \
        break;
    switch (rows) {
        DPA_ROWS(6)
        DPA_ROWS(8)
        DPA_ROWS(10)
    default:
        k_fused_pair_batch<8, Access<true>><<<blocks, 256, 0, stream>>>(instances, slots, owners, n_tiles, injections,
                                                                    xcd_group);
        break;
    }
#undef DPA_ROWS
}

void launch_fused_pair_batch(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                             uint32_t n_tiles, const InjDesc *injections, int rows, bool nontemporal, bool buffer,
                             uint32_t xcd_group, hipStream_t stream) {
    if (n_tiles == 0) return;
    if (buffer)
        launch_pair_batch_as<BufferAccess>(instances, slots, owners, n_tiles, injections, rows, nontemporal, xcd_group,
                                           stream);
    else
        launch_pair_batch_as<Pointer64>(instances, slots, owners, n_tiles, injections, rows, nontemporal, xcd_group,
                                        stream);
}

}  // namespace dpa
