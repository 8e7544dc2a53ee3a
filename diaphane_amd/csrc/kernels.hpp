// Device-side data structures and kernels for the Yee step on gfx950.
//
// Layout in HBM (one allocation each):
//   E[2], H[2]  3·N f32, component-major, x fastest (ping-pong pair)
//   material    N u8 (index into a ≤256-entry coefficient table)
//   axis table  4 sections × (nx+ny+nz) f32:
//               [primary gains x|y|z][dual gains x|y|z]
//               [integer absorber x|y|z][half absorber x|y|z]
//   intensity   N f64
//
// Arithmetic contract: every cell update is the reference's expression in
// the reference's order (src/cpu.rs:319-331, 418-430), compiled with
// -ffp-contract=off and IEEE division, so results are bit-identical to the
// CPU sweep.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dpa {

// One injection as the kernels consume it: a box, a component, a value
// already multiplied by the direction of time, and an optional offset into a
// host-computed apodization-weight table laid out over the box, x fastest.
struct alignas(16) InjDesc {
    uint32_t origin[3];
    uint32_t extent[3];
    uint32_t component;
    int32_t weight_offset;  // < 0: weight 1
    float value;
    uint32_t pad[3];
};

// Sections of the per-axis table.
enum : int { kGainH = 0, kGainE = 1, kLossInt = 2, kLossHalf = 3 };

struct StepArgs {
    const float *e_in;
    const float *h_in;
    float *e_out;
    float *h_out;
    const uint8_t *material;
    const float4 *coefficients;  // (eg, el, mg, ml) per material
    const float *axis_table;     // see above
    const float *weights;        // apodization weights
    const InjDesc *injections;   // magnetic ones first, then electric ones
    double *intensity;           // null when not accumulating
    uint32_t n_magnetic;
    uint32_t n_electric;
    uint32_t material_count;
    uint32_t nx, ny, nz;
    uint64_t cells;
    // fused-kernel tiling
    uint32_t tiles_x, tiles_y, tiles_z, depth;
    uint32_t tiles_total;
    uint32_t xcd_remap;  // 0, or the run length in blocks per XCD
    // the second step of a two-step launch (k_fused_pair): its injections,
    // laid out like the first step's
    const InjDesc *injections2;
    uint32_t n_magnetic2;
    uint32_t n_electric2;
    uint32_t single_step;  // a two-step launch that advances this grid one step
};

// One grid of a batch, as the batch kernel sees it (device memory).
struct alignas(16) InstanceDesc {
    float *e[2];  // ping-pong pair
    float *h[2];
    const uint8_t *material;
    const float4 *coefficients;
    const float *axis_table;
    const float *weights;
    double *intensity;
    uint32_t nx, ny, nz, material_count;
    uint64_t cells;
    uint32_t tiles_x, tiles_y, tiles_z, depth;
    uint32_t tiles_total, tile_base, pad0, pad1;
};

// What one instance does in one batch launch.
struct alignas(16) StepSlot {
    uint32_t active;      // steps this launch
    uint32_t parity;      // which buffer of the pair holds the state
    uint32_t inj_offset;  // into the launch's injection list
    uint32_t n_magnetic;
    uint32_t n_electric;
    uint32_t accumulate;  // add |E|² into the instance's intensity
    // two-step launches: the second step's injections, and 1 when this
    // instance takes only one step (its last, in an odd count)
    uint32_t inj_offset2;
    uint32_t n_magnetic2;
    uint32_t n_electric2;
    uint32_t single;
    uint32_t pad[2];
};

// Tile geometry of the fused kernel: one wave of 64 lanes covers 62 owned
// columns plus one halo column on each side.
constexpr int kLanes = 64;
constexpr int kOwnedX = kLanes - 2;

// 62-wide tiles whose halo columns are ordinary lanes (the first fused form).
void launch_fused(const StepArgs &args, int rows, hipStream_t stream);
// 64-aligned tiles; halo columns fetched by lanes 0 and 63.
constexpr int kAlignedX = 64;
// `buffer`: address fields through buffer resources (32-bit lane offsets,
// component offsets in the scalar offset) when the grid allows it.
void launch_fused_aligned(const StepArgs &args, int rows, bool nontemporal, bool buffer, hipStream_t stream);
bool buffer_addressing_fits(uint64_t cells, uint32_t nx, uint32_t ny);
// Two steps per launch: 64-lane waves owning 60 columns (two halo lanes on
// each side) and rows − 4 of `rows` register rows (6, 8 or 10).
constexpr int kPairOwnedX = kLanes - 4;
// Row tiles of stacked two-step tiles of `rows` rows in all (waves × rows per
// wave). Tile t's first row sits at y = (rows − 4)·t. Every half-step of a
// stacked tile runs on all its rows, so tile 0 also owns its first two rows
// (below them lies the low wall), and the last tile every row from 2 inside
// the grid, the count leaving its top row past the high wall.
inline uint32_t stack_tiles_y(uint32_t ny, uint32_t rows) {
    return ny < rows ? 1u : 1u + (ny - rows + 1 + (rows - 4) - 1) / (rows - 4);
}
void launch_fused_pair(const StepArgs &args, int rows, bool nontemporal, bool buffer, hipStream_t stream);
// Many grids, two steps per launch: one wave per tile of the packed table
// (tiles cut for the two-step kernel). `xcd_group` > 0 gives each XCD runs of
// that many consecutive workgroups; 0 keeps the hardware's round robin.
void launch_fused_pair_batch(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                             uint32_t n_tiles, const InjDesc *injections, int rows, bool nontemporal, bool buffer,
                             uint32_t xcd_group, hipStream_t stream);
// Two steps per launch with `waves` (2 or 4) waves of a workgroup stacked
// along y on one tile of waves·rows register rows (rows 4 or 6), of which
// waves·rows − 4 are owned; neighbouring waves trade edge rows through LDS.
// One workgroup per tile.
void launch_fused_stack(const StepArgs &args, int rows, int waves, bool buffer, hipStream_t stream);
void launch_fused_stack_batch(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                              uint32_t n_tiles, const InjDesc *injections, int rows, int waves, bool buffer,
                              uint32_t xcd_group, hipStream_t stream);
// One launch advancing every active instance of a batch by one step.
// `blocks` is the persistent grid size (workgroups of 4 waves); with
// `direct` the launch instead has one wave per tile.
// `buffer` selects buffer-resource addressing; the host guarantees that every
// instance's offsets fit it.
void launch_fused_batch(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                        uint32_t n_tiles, const InjDesc *injections, int rows, bool nontemporal, bool buffer,
                        bool direct, uint32_t blocks, hipStream_t stream);
// Workgroups of the batch kernel that fit on the device at once.
uint32_t batch_resident_blocks(int rows, bool nontemporal);
void launch_update_h(const StepArgs &args, int direction, hipStream_t stream);
void launch_update_e(const StepArgs &args, int direction, hipStream_t stream);
void launch_inject(float *field, uint64_t cells, uint32_t nx, uint32_t ny, const InjDesc &desc,
                   const InjDesc *device_desc, const float *weights, hipStream_t stream);
void launch_intensity(const float *e, double *intensity, uint64_t cells, hipStream_t stream);
// out[i] = the f32 rounding of sum[i] (0 where sum is null).
void launch_intensity_f32(const double *sum, float *out, uint64_t cells, hipStream_t stream);
// f64 energy partials, one pair per block; returns the number of blocks.
uint32_t launch_energy(const float *e, const float *h, const uint8_t *material,
                       const float2 *eps_mu, const float *widths, uint32_t nx, uint32_t ny,
                       uint32_t nz, double *partials, hipStream_t stream);
uint32_t energy_blocks(uint32_t nx, uint32_t ny, uint32_t nz);
// The viewer's auto-range maxima: per block eight floats into `partials`
// (peak_blocks(cells) × 8); returns the number of blocks. `intensity` may be
// null; `steps` is the accumulated step count (0 counts as 1).
uint32_t peak_blocks(uint64_t cells);
uint32_t launch_measure_peak(const float *e, const float *h, const double *intensity, uint64_t steps, uint64_t cells,
                             float *partials, hipStream_t stream);
// Streams the step's access mix (E, H, material in; E, H out) with no
// stencil: the practical bandwidth ceiling for the step's byte count.
void launch_stream_probe(const float *e, const float *h, const uint8_t *material, float *e_out, float *h_out,
                         uint64_t cells, hipStream_t stream);
// The fused kernel's tile access pattern (rows 4, 6 or 8; the tiling in
// `args`) with no stencil: the ceiling for that tiling's memory traffic.
void launch_tile_probe(const StepArgs &args, int rows, hipStream_t stream);

}  // namespace dpa
