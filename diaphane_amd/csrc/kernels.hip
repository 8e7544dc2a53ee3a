// Yee-step kernels for gfx950. See kernels.hpp for the layout and the
// arithmetic contract.
#include "kernels.hpp"

#include <algorithm>

namespace dpa {

namespace {

#include "device_common.inc"


// H ← ((1−m)·H − g·curl)/(1+m) forward (src/cpu.rs:328-330) and its exact
// inverse ((1+m)·H + g·curl)/(1−m) (src/cpu.rs:525-527). With m = 0 the
// forward form is H − g·curl exactly ((1−0)·H = H and x/1 = x in IEEE
// arithmetic, NaN and signed zero included), which skips the divide.
template <int Direction>
__device__ __forceinline__ float magnetic_update(float h, float loss, float gain, float curl) {
    if (Direction > 0) {
        if (loss == 0.0f) return h - gain * curl;
        return ((1.0f - loss) * h - gain * curl) / (1.0f + loss);
    }
    return ((1.0f + loss) * h + gain * curl) / (1.0f - loss);
}

// E ← ((1−l)·E + g·curl)/(1+l) (src/cpu.rs:427-429) and its inverse
// ((1+l)·E − g·curl)/(1−l) (src/cpu.rs:480-482).
template <int Direction>
__device__ __forceinline__ float electric_update(float e, float loss, float gain, float curl) {
    if (Direction > 0) {
        if (loss == 0.0f) return e + gain * curl;
        return ((1.0f - loss) * e + gain * curl) / (1.0f + loss);
    }
    return ((1.0f + loss) * e - gain * curl) / (1.0f - loss);
}

// Soft injections at one cell, in list order: F += value · weight
// (src/cpu.rs:606-633). Descriptor reads are wave-uniform.
__device__ __forceinline__ float inject(float f, uint32_t component, uint32_t x, uint32_t y,
                                        uint32_t z, const InjDesc *__restrict__ list,
                                        uint32_t count, const float *__restrict__ weights) {
    for (uint32_t n = 0; n < count; ++n) {
        const InjDesc &q = list[n];
        const uint32_t dx = x - q.origin[0], dy = y - q.origin[1], dz = z - q.origin[2];
        if (q.component == component && dx < q.extent[0] && dy < q.extent[1] && dz < q.extent[2]) {
            float w = 1.0f;
            if (q.weight_offset >= 0)
                w = weights[q.weight_offset + (dz * q.extent[1] + dy) * q.extent[0] + dx];
            f = f + q.value * w;
        }
    }
    return f;
}



// ---------------------------------------------------------------------------
// Fused step: H, H-injections, E, E-injections and intensity in one pass.
//
// One wave owns a tile of 62 (x) × (R−2) (y) columns and marches a z-slab of
// `depth` planes. Lane l sits at x = 62·tx − 1 + l and row r at
// y = (R−2)·ty − 1 + r, so lane 0 / row 0 are the low halo and lane 63 /
// row R−1 the high halo. Per plane k the wave
//   1. loads E(k+1), H(k) and the material of plane k (E(k) is carried in
//      registers from the previous plane),
//   2. computes H_new(k) on lanes 0..62 × rows 0..R−2 — the owned cells plus
//      the low halo that E needs (halo values are recomputed, not shared),
//   3. adds magnetic injections,
//   4. computes E_new(k) on the owned cells using H_new(k), its x−1
//      neighbour (lane shuffle), y−1 (register) and H_new(k−1) (register),
//   5. adds electric injections and the intensity term, and stores.
// A slab starting at zb > 0 first recomputes H_new(zb−1) as its z halo.
// E and H are read from one buffer of the ping-pong pair and written to the
// other, so halo reads never see a neighbour's updated values.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void k_fused(StepArgs p) {
    static_assert(R >= 3, "a tile needs a halo row on each side");
    __shared__ float4 s_coeff[256];
    for (uint32_t m = threadIdx.x; m < p.material_count; m += blockDim.x) s_coeff[m] = p.coefficients[m];
    __syncthreads();

    uint32_t block = blockIdx.x;
    if (p.xcd_remap) block = xcd_grouped(block, gridDim.x, p.xcd_remap);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(block * (blockDim.x / kLanes) + threadIdx.x / kLanes);
    if (tile >= p.tiles_total) return;
    const int lane = threadIdx.x & (kLanes - 1);

    const uint32_t nx = p.nx, ny = p.ny, nz = p.nz;
    const uint32_t span = nx + ny + nz;
    const uint64_t plane = uint64_t(nx) * ny;
    const uint64_t cells = p.cells;
    const float *__restrict__ table = p.axis_table;
    const float *__restrict__ e_in = p.e_in;
    const float *__restrict__ h_in = p.h_in;
    float *__restrict__ e_out = p.e_out;
    float *__restrict__ h_out = p.h_out;
    const uint8_t *__restrict__ material = p.material;
    const InjDesc *__restrict__ magnetic_list = p.injections;
    const InjDesc *__restrict__ electric_list = p.injections + p.n_magnetic;
    const float *__restrict__ weights = p.weights;

    const uint32_t tx = tile % p.tiles_x;
    const uint32_t rest = tile / p.tiles_x;
    const uint32_t ty = rest % p.tiles_y;
    const uint32_t tz = rest / p.tiles_y;
    const int x = int(tx * kOwnedX) - 1 + lane;
    const int y0 = int(ty * (R - 2)) - 1;
    const uint32_t zb = tz * p.depth;
    const uint32_t ze = min(zb + p.depth, nz);
    const uint32_t zs = zb > 0 ? zb - 1 : 0;

    const bool x_in = x >= 0 && x < int(nx);
    const bool x_owned = x_in && lane >= 1 && lane <= kOwnedX;
    float gain_h_x = 0.0f, gain_e_x = 0.0f, loss_int_x = 0.0f, loss_half_x = 0.0f;
    if (x_in) {
        gain_h_x = axis_value(table, span, kGainH, x);
        gain_e_x = axis_value(table, span, kGainE, x);
        loss_int_x = axis_value(table, span, kLossInt, x);
        loss_half_x = axis_value(table, span, kLossHalf, x);
    }
    bool y_in[R];
    float gain_h_y[R], gain_e_y[R], loss_int_y[R], loss_half_y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y0 + r;
        y_in[r] = y >= 0 && y < int(ny);
        const uint32_t yc = y_in[r] ? uint32_t(y) : 0;
        gain_h_y[r] = axis_value(table, span, kGainH, nx + yc);
        gain_e_y[r] = axis_value(table, span, kGainE, nx + yc);
        loss_int_y[r] = axis_value(table, span, kLossInt, nx + yc);
        loss_half_y[r] = axis_value(table, span, kLossHalf, nx + yc);
    }

    // Column offset of (x, y0 + r) within a plane; only dereferenced when in.
    auto column = [&](int r) -> int64_t { return int64_t(y0 + r) * nx + x; };

    float ex[R], ey[R], ez[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        ex[r] = ey[r] = ez[r] = 0.0f;
        if (x_in && y_in[r]) {
            const uint64_t i = uint64_t(zs) * plane + column(r);
            ex[r] = e_in[i];
            ey[r] = e_in[cells + i];
            ez[r] = e_in[2 * cells + i];
        }
    }
    float hx_prev[R], hy_prev[R];
#pragma unroll
    for (int r = 0; r < R; ++r) hx_prev[r] = hy_prev[r] = 0.0f;

    for (uint32_t k = zs; k < ze; ++k) {
        const uint32_t kz = nx + ny + k;
        const float gain_h_z = axis_value(table, span, kGainH, kz);
        const float gain_e_z = axis_value(table, span, kGainE, kz);
        const float loss_int_z = axis_value(table, span, kLossInt, kz);
        const float loss_half_z = axis_value(table, span, kLossHalf, kz);
        const bool next_in = k + 1 < nz;

        float nex[R], ney[R], nez[R];
        float hx[R - 1], hy[R - 1], hz[R - 1];
        uint32_t mat[R - 1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            nex[r] = ney[r] = nez[r] = 0.0f;
            const bool in = x_in && y_in[r];
            if (in && next_in) {
                const uint64_t i = uint64_t(k + 1) * plane + column(r);
                nex[r] = e_in[i];
                ney[r] = e_in[cells + i];
                nez[r] = e_in[2 * cells + i];
            }
            if (r < R - 1) {
                hx[r] = hy[r] = hz[r] = 0.0f;
                mat[r] = 0;
                if (in) {
                    const uint64_t i = uint64_t(k) * plane + column(r);
                    hx[r] = h_in[i];
                    hy[r] = h_in[cells + i];
                    hz[r] = h_in[2 * cells + i];
                    mat[r] = material[i];
                }
            }
        }

        // H_new(k): forward differences; E past a high wall reads as the
        // implicit zero the masked loads already supply (src/cpu.rs:335-374).
#pragma unroll
        for (int r = 0; r < R - 1; ++r) {
            const float ez_xp = __shfl_down(ez[r], 1, kLanes);
            const float ey_xp = __shfl_down(ey[r], 1, kLanes);
            const float4 c = s_coeff[mat[r]];
            const float mg = c.z, ml = c.w;
            {
                const float curl = gain_h_y[r] * (ez[r + 1] - ez[r]) - gain_h_z * (ney[r] - ey[r]);
                const float loss = (ml + (loss_half_y[r] + loss_half_z)) + loss_int_x;
                hx[r] = magnetic_update<1>(hx[r], loss, mg, curl);
            }
            {
                const float curl = gain_h_z * (nex[r] - ex[r]) - gain_h_x * (ez_xp - ez[r]);
                const float loss = (ml + (loss_int_y[r] + loss_half_z)) + loss_half_x;
                hy[r] = magnetic_update<1>(hy[r], loss, mg, curl);
            }
            {
                const float curl = gain_h_x * (ey_xp - ey[r]) - gain_h_y[r] * (ex[r + 1] - ex[r]);
                const float loss = (ml + (loss_half_y[r] + loss_int_z)) + loss_half_x;
                hz[r] = magnetic_update<1>(hz[r], loss, mg, curl);
            }
            if (p.n_magnetic) {
                const uint32_t yy = uint32_t(y0 + r), xx = uint32_t(x);
                hx[r] = inject(hx[r], 0, xx, yy, k, magnetic_list, p.n_magnetic, weights);
                hy[r] = inject(hy[r], 1, xx, yy, k, magnetic_list, p.n_magnetic, weights);
                hz[r] = inject(hz[r], 2, xx, yy, k, magnetic_list, p.n_magnetic, weights);
            }
        }

        if (k >= zb) {
#pragma unroll
            for (int r = 1; r < R - 1; ++r) {
                const float hy_xm = __shfl_up(hy[r], 1, kLanes);
                const float hz_xm = __shfl_up(hz[r], 1, kLanes);
                const int y = y0 + r;
                const float4 c = s_coeff[mat[r]];
                const float eg = c.x, el = c.y;
                float ex_new = ex[r], ey_new = ey[r], ez_new = ez[r];
                // Index 0 along b or c is never updated: the low-wall PEC
                // (src/cpu.rs:392-401).
                if (y > 0 && k > 0) {
                    const float curl = gain_e_y[r] * (hz[r] - hz[r - 1]) - gain_e_z * (hy[r] - hy_prev[r]);
                    const float loss = (el + (loss_int_y[r] + loss_int_z)) + loss_half_x;
                    ex_new = electric_update<1>(ex[r], loss, eg, curl);
                }
                if (k > 0 && x > 0) {
                    const float curl = gain_e_z * (hx[r] - hx_prev[r]) - gain_e_x * (hz[r] - hz_xm);
                    const float loss = (el + (loss_half_y[r] + loss_int_z)) + loss_int_x;
                    ey_new = electric_update<1>(ey[r], loss, eg, curl);
                }
                if (x > 0 && y > 0) {
                    const float curl = gain_e_x * (hy[r] - hy_xm) - gain_e_y[r] * (hx[r] - hx[r - 1]);
                    const float loss = (el + (loss_int_y[r] + loss_half_z)) + loss_int_x;
                    ez_new = electric_update<1>(ez[r], loss, eg, curl);
                }
                if (p.n_electric) {
                    const uint32_t yy = uint32_t(y), xx = uint32_t(x);
                    ex_new = inject(ex_new, 0, xx, yy, k, electric_list, p.n_electric, weights);
                    ey_new = inject(ey_new, 1, xx, yy, k, electric_list, p.n_electric, weights);
                    ez_new = inject(ez_new, 2, xx, yy, k, electric_list, p.n_electric, weights);
                }
                if (x_owned && y_in[r]) {
                    const uint64_t i = uint64_t(k) * plane + column(r);
                    h_out[i] = hx[r];
                    h_out[cells + i] = hy[r];
                    h_out[2 * cells + i] = hz[r];
                    e_out[i] = ex_new;
                    e_out[cells + i] = ey_new;
                    e_out[2 * cells + i] = ez_new;
                    if (p.intensity) p.intensity[i] += double(intensity_term(ex_new, ey_new, ez_new));
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R - 1; ++r) {
            hx_prev[r] = hx[r];
            hy_prev[r] = hy[r];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ex[r] = nex[r];
            ey[r] = ney[r];
            ez[r] = nez[r];
        }
    }
}

#include "fused_access.inc"
#include "fused_aligned.inc"

// ---------------------------------------------------------------------------
// Two-pass path: the reference's dispatch shape (src/shaders/fdtd.wgsl:139-214),
// one thread per cell, updating in place. Used for reversal and as an
// independent cross-check of the fused kernel.
// ---------------------------------------------------------------------------
template <int Direction>
__global__ __launch_bounds__(256) void k_update_h(StepArgs p) {
    const uint32_t x = blockIdx.x * 64 + threadIdx.x;
    const uint32_t y = blockIdx.y * 4 + threadIdx.y;
    const uint32_t z = blockIdx.z;
    const uint32_t nx = p.nx, ny = p.ny, nz = p.nz;
    if (x >= nx || y >= ny) return;
    const uint32_t span = nx + ny + nz;
    const uint64_t sy = nx, sz = uint64_t(nx) * ny, n = p.cells;
    const uint64_t i = uint64_t(z) * sz + uint64_t(y) * nx + x;
    const float *__restrict__ t = p.axis_table;
    const float *ex = p.e_in, *ey = p.e_in + n, *ez = p.e_in + 2 * n;
    float *hx = p.h_out, *hy = p.h_out + n, *hz = p.h_out + 2 * n;
    const float4 c = p.coefficients[p.material[i]];
    const float mg = c.z, ml = c.w;
    const float gx = axis_value(t, span, kGainH, x), gy = axis_value(t, span, kGainH, nx + y),
                gz = axis_value(t, span, kGainH, nx + ny + z);
    const float pix = axis_value(t, span, kLossInt, x), piy = axis_value(t, span, kLossInt, nx + y),
                piz = axis_value(t, span, kLossInt, nx + ny + z);
    const float phx = axis_value(t, span, kLossHalf, x), phy = axis_value(t, span, kLossHalf, nx + y),
                phz = axis_value(t, span, kLossHalf, nx + ny + z);
    const float exc = ex[i], eyc = ey[i], ezc = ez[i];
    const float ex_yp = y + 1 < ny ? ex[i + sy] : 0.0f, ex_zp = z + 1 < nz ? ex[i + sz] : 0.0f;
    const float ey_xp = x + 1 < nx ? ey[i + 1] : 0.0f, ey_zp = z + 1 < nz ? ey[i + sz] : 0.0f;
    const float ez_xp = x + 1 < nx ? ez[i + 1] : 0.0f, ez_yp = y + 1 < ny ? ez[i + sy] : 0.0f;
    {
        const float curl = gy * (ez_yp - ezc) - gz * (ey_zp - eyc);
        hx[i] = magnetic_update<Direction>(hx[i], (ml + (phy + phz)) + pix, mg, curl);
    }
    {
        const float curl = gz * (ex_zp - exc) - gx * (ez_xp - ezc);
        hy[i] = magnetic_update<Direction>(hy[i], (ml + (piy + phz)) + phx, mg, curl);
    }
    {
        const float curl = gx * (ey_xp - eyc) - gy * (ex_yp - exc);
        hz[i] = magnetic_update<Direction>(hz[i], (ml + (phy + piz)) + phx, mg, curl);
    }
}

template <int Direction>
__global__ __launch_bounds__(256) void k_update_e(StepArgs p) {
    const uint32_t x = blockIdx.x * 64 + threadIdx.x;
    const uint32_t y = blockIdx.y * 4 + threadIdx.y;
    const uint32_t z = blockIdx.z;
    const uint32_t nx = p.nx, ny = p.ny, nz = p.nz;
    if (x >= nx || y >= ny) return;
    const uint32_t span = nx + ny + nz;
    const uint64_t sy = nx, sz = uint64_t(nx) * ny, n = p.cells;
    const uint64_t i = uint64_t(z) * sz + uint64_t(y) * nx + x;
    const float *__restrict__ t = p.axis_table;
    const float *hx = p.h_in, *hy = p.h_in + n, *hz = p.h_in + 2 * n;
    float *ex = p.e_out, *ey = p.e_out + n, *ez = p.e_out + 2 * n;
    const float4 c = p.coefficients[p.material[i]];
    const float eg = c.x, el = c.y;
    const float gx = axis_value(t, span, kGainE, x), gy = axis_value(t, span, kGainE, nx + y),
                gz = axis_value(t, span, kGainE, nx + ny + z);
    const float pix = axis_value(t, span, kLossInt, x), piy = axis_value(t, span, kLossInt, nx + y),
                piz = axis_value(t, span, kLossInt, nx + ny + z);
    const float phx = axis_value(t, span, kLossHalf, x), phy = axis_value(t, span, kLossHalf, nx + y),
                phz = axis_value(t, span, kLossHalf, nx + ny + z);
    if (y > 0 && z > 0) {
        const float curl = gy * (hz[i] - hz[i - sy]) - gz * (hy[i] - hy[i - sz]);
        ex[i] = electric_update<Direction>(ex[i], (el + (piy + piz)) + phx, eg, curl);
    }
    if (z > 0 && x > 0) {
        const float curl = gz * (hx[i] - hx[i - sz]) - gx * (hz[i] - hz[i - 1]);
        ey[i] = electric_update<Direction>(ey[i], (el + (phy + piz)) + pix, eg, curl);
    }
    if (x > 0 && y > 0) {
        const float curl = gx * (hy[i] - hy[i - 1]) - gy * (hx[i] - hx[i - sy]);
        ez[i] = electric_update<Direction>(ez[i], (el + (piy + phz)) + pix, eg, curl);
    }
}

// One injection over its box (src/shaders/fdtd.wgsl:237-272), weights taken
// from the host-computed table rather than evaluated on the device.
__global__ __launch_bounds__(256) void k_inject(float *field, uint32_t nx, uint32_t ny, InjDesc d,
                                                const float *weights) {
    const uint32_t dx = blockIdx.x * 256 + threadIdx.x;
    const uint32_t dy = blockIdx.y, dz = blockIdx.z;
    if (dx >= d.extent[0]) return;
    const uint64_t i = (uint64_t(d.origin[2] + dz) * ny + (d.origin[1] + dy)) * nx + d.origin[0] + dx;
    float w = 1.0f;
    if (d.weight_offset >= 0) w = weights[d.weight_offset + (dz * d.extent[1] + dy) * d.extent[0] + dx];
    field[i] = field[i] + d.value * w;
}

// The f64 intensity sum rounded to f32, for a renderer that binds f32.
__global__ __launch_bounds__(256) void k_intensity_f32(const double *sum, float *out, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = sum ? float(sum[i]) : 0.0f;
}

__global__ __launch_bounds__(256) void k_intensity(const float *e, double *intensity, uint64_t n) {
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    intensity[i] += double(intensity_term(e[i], e[n + i], e[2 * n + i]));
}

// f64 energy with per-component volumes (src/cpu.rs:654-687). Each block
// writes one (electric, magnetic) partial; the host adds them in order, so
// the result is deterministic.
constexpr uint32_t kEnergyThreads = 256;

// The viewer's auto-range reduction (examples/common/render.rs:382-419,
// shaders/volume.wgsl:122-132, 376-378, 473-493), eight maxima per block:
//   [0]    max over cells of max(½|E|², ½|H|²), each sum ((0+x²)+y²)+z² in f32
//   [1..3] max |E_x|, |E_y|, |E_z|;  [4..6] max |H_x|, |H_y|, |H_z|
//   [7]    max of the intensity sum (rounded to f32) / f32(steps), 0 if none.
// Maxima are exact, so the result does not depend on the block split; the
// host reduces the per-block partials.
constexpr uint32_t kPeakThreads = 256;
constexpr uint32_t kPeakSlots = 8;

__global__ __launch_bounds__(kPeakThreads) void k_measure_peak(const float *e, const float *h,
                                                               const double *intensity, float steps,
                                                               uint64_t n, float *partials) {
    float m[kPeakSlots] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (uint64_t i = uint64_t(blockIdx.x) * kPeakThreads + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * kPeakThreads) {
        float se = 0.f, sh = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float ev = e[c * n + i], hv = h[c * n + i];
            se = se + ev * ev;
            sh = sh + hv * hv;
            m[1 + c] = fmaxf(m[1 + c], fabsf(ev));
            m[4 + c] = fmaxf(m[4 + c], fabsf(hv));
        }
        m[0] = fmaxf(m[0], fmaxf(0.5f * se, 0.5f * sh));
        if (intensity) m[7] = fmaxf(m[7], float(intensity[i]) / steps);
    }
    __shared__ float s_peak[kPeakSlots][kPeakThreads];
#pragma unroll
    for (uint32_t k = 0; k < kPeakSlots; ++k) s_peak[k][threadIdx.x] = m[k];
    __syncthreads();
    for (uint32_t stride = kPeakThreads / 2; stride > 0; stride /= 2) {
        if (threadIdx.x < stride) {
#pragma unroll
            for (uint32_t k = 0; k < kPeakSlots; ++k)
                s_peak[k][threadIdx.x] = fmaxf(s_peak[k][threadIdx.x], s_peak[k][threadIdx.x + stride]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kPeakSlots) partials[blockIdx.x * kPeakSlots + threadIdx.x] = s_peak[threadIdx.x][0];
}

__global__ __launch_bounds__(kEnergyThreads) void k_energy(const float *e, const float *h,
                                                           const uint8_t *material,
                                                           const float2 *eps_mu,
                                                           const float *widths, uint32_t nx,
                                                           uint32_t ny, uint32_t nz,
                                                           double *partials) {
    const uint64_t n = uint64_t(nx) * ny * nz;
    const uint32_t span = nx + ny + nz;
    const float *primary = widths, *dual = widths + span;
    const uint32_t base[3] = {0, nx, nx + ny};
    double ue = 0.0, uh = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * kEnergyThreads + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * kEnergyThreads) {
        const uint32_t coord[3] = {uint32_t(i % nx), uint32_t((i / nx) % ny), uint32_t(i / (uint64_t(nx) * ny))};
        const float2 m = eps_mu[material[i]];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int b = (a + 1) % 3, c = (a + 2) % 3;
            const float ev = e[a * n + i], hv = h[a * n + i];
            const double electric_volume = double(primary[base[a] + coord[a]]) *
                                           double(dual[base[b] + coord[b]]) *
                                           double(dual[base[c] + coord[c]]);
            const double magnetic_volume = double(dual[base[a] + coord[a]]) *
                                           double(primary[base[b] + coord[b]]) *
                                           double(primary[base[c] + coord[c]]);
            ue += 0.5 * electric_volume * double(m.x * ev * ev);
            uh += 0.5 * magnetic_volume * double(m.y * hv * hv);
        }
    }
    __shared__ double s_e[kEnergyThreads], s_h[kEnergyThreads];
    s_e[threadIdx.x] = ue;
    s_h[threadIdx.x] = uh;
    __syncthreads();
    for (uint32_t stride = kEnergyThreads / 2; stride > 0; stride /= 2) {
        if (threadIdx.x < stride) {
            s_e[threadIdx.x] += s_e[threadIdx.x + stride];
            s_h[threadIdx.x] += s_h[threadIdx.x + stride];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = s_e[0];
        partials[2 * blockIdx.x + 1] = s_h[0];
    }
}

// The Yee step's HBM access mix with no stencil and no arithmetic: six f32
// streams and one u8 stream read, six f32 streams written, 16 bytes per lane
// per stream. Measures the bandwidth this access pattern can sustain.
__device__ __forceinline__ void store_stream(float4 v, float4 *p) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const v4 x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<v4 *>(p));
}

// One float4 of each of the six field streams (and 4 material bytes) per
// thread, stores non-temporal. Each load's address depends on the previous
// load's value through `zero` (always 0), so a wave has one load in flight at
// a time; launched at four workgroups per CU. Measured on MI355X this is the
// fastest streaming form of the mix (about 6.2 TB/s): issuing all seven
// loads at once, or running more waves per CU, congests the memory system
// and streams about 12% slower.
__global__ __launch_bounds__(256) void k_stream_probe(const float4 *__restrict__ e, const float4 *__restrict__ h,
                                                      const uint32_t *__restrict__ material,
                                                      float4 *__restrict__ e_out, float4 *__restrict__ h_out,
                                                      uint64_t quads, uint32_t zero) {
    const uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= quads) return;
    float4 a[3], b[3];
    uint64_t dep = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a[c] = e[c * quads + q + dep];
        dep = __float_as_uint(a[c].x) & zero;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b[c] = h[c * quads + q + dep];
        dep = __float_as_uint(b[c].x) & zero;
    }
    const float bias = float(material[q + dep] & 1u);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        a[c].x += bias;
        b[c].x += bias;
        store_stream(a[c], &e_out[c * quads + q]);
        store_stream(b[c], &h_out[c * quads + q]);
    }
}

// The fused kernel's access pattern with no stencil: the same 64-aligned
// tiles of R rows marching a z-slab, loading E at plane k+1, H and the
// material at plane k for every row and the x-halo cells (lanes 0 and 63),
// and storing E and H of the owned rows non-temporally. Its bandwidth is the
// ceiling for that tiling's memory traffic; the stencil kernel can approach
// it, not pass it.
template <int R>
__global__ __launch_bounds__(256) void k_tile_probe(StepArgs p) {
    const uint32_t tile = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / kLanes);
    if (tile >= p.tiles_total) return;
    const int lane = threadIdx.x & (kLanes - 1);
    const int nx = int(p.nx), ny = int(p.ny), nz = int(p.nz);
    const uint64_t plane = uint64_t(p.nx) * p.ny, cells = p.cells;
    const uint32_t tx = tile % p.tiles_x, rest = tile / p.tiles_x;
    const uint32_t ty = rest % p.tiles_y, tz = rest / p.tiles_y;
    const int x0 = int(tx) * kLanes, x = x0 + lane;
    const int xc = x < nx ? x : nx - 1;
    const bool is_lo = lane == 0, is_hi = lane == kLanes - 1;
    const int xh = is_lo ? (x0 > 0 ? x0 - 1 : 0) : (is_hi ? (x0 + kLanes < nx ? x0 + kLanes : xc) : xc);
    const int y0 = int(ty) * (R - 2) - 1;
    const int zb = int(tz * p.depth);
    const int ze = min(zb + int(p.depth), nz);
    const int zs = zb > 0 ? zb - 1 : 0;
    uint64_t own[R], halo[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int y = y0 + r;
        const int yc = y < 0 ? 0 : (y < ny ? y : ny - 1);
        own[r] = uint64_t(yc) * uint64_t(nx) + uint64_t(xc);
        halo[r] = uint64_t(yc) * uint64_t(nx) + uint64_t(xh);
    }
    for (int k = zs; k < ze; ++k) {
        const uint64_t pk = uint64_t(k) * plane;
        const uint64_t pn = uint64_t(k + 1 < nz ? k + 1 : k) * plane;
        float ev[3][R], hv[3][R], eg[3][R], hg[3][R];
        uint32_t mv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                ev[c][r] = p.e_in[c * cells + pn + own[r]];
                hv[c][r] = p.h_in[c * cells + pk + own[r]];
                eg[c][r] = p.e_in[c * cells + pn + halo[r]];
                hg[c][r] = p.h_in[c * cells + pk + halo[r]];
            }
            mv[r] = p.material[pk + own[r]];
        }
        if (k < zb) continue;
#pragma unroll
        for (int r = 1; r < R - 1; ++r) {
            if (y0 + r >= ny || x >= nx) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float ve = ev[c][r] + hv[c][r - 1] + float(mv[r] & 1u) + (is_hi ? eg[c][r] : 0.0f);
                const float vh = hv[c][r] + ev[c][r + 1] + (is_lo ? hg[c][r] : 0.0f);
                __builtin_nontemporal_store(ve, p.e_out + c * cells + pk + own[r]);
                __builtin_nontemporal_store(vh, p.h_out + c * cells + pk + own[r]);
            }
        }
    }
}

}  // namespace

void launch_stream_probe(const float *e, const float *h, const uint8_t *material, float *e_out, float *h_out,
                         uint64_t cells, hipStream_t stream) {
    const uint64_t quads = cells / 4;
    const uint32_t blocks = uint32_t((quads + 255) / 256);
    // 40 KiB of (unused) dynamic LDS per workgroup: four workgroups per CU.
    k_stream_probe<<<blocks, 256, 40 * 1024, stream>>>(
        reinterpret_cast<const float4 *>(e), reinterpret_cast<const float4 *>(h),
        reinterpret_cast<const uint32_t *>(material), reinterpret_cast<float4 *>(e_out),
        reinterpret_cast<float4 *>(h_out), quads, 0u);
}

void launch_tile_probe(const StepArgs &args, int rows, hipStream_t stream) {
    const uint32_t blocks = (args.tiles_total + 3) / 4;
    switch (rows) {
    case 6: k_tile_probe<6><<<blocks, 256, 0, stream>>>(args); break;
    case 8: k_tile_probe<8><<<blocks, 256, 0, stream>>>(args); break;
    default: k_tile_probe<4><<<blocks, 256, 0, stream>>>(args); break;
    }
}

void launch_fused(const StepArgs &args, int rows, hipStream_t stream) {
    const uint32_t blocks = (args.tiles_total + 3) / 4;
    switch (rows) {
#define DPA_ROWS(R) \
    case R: k_fused<R><<<blocks, 256, 0, stream>>>(args); break;
        DPA_ROWS(4)
        DPA_ROWS(6)
        DPA_ROWS(8)
        DPA_ROWS(10)
        DPA_ROWS(12)
#undef DPA_ROWS
    default:
        k_fused<8><<<blocks, 256, 0, stream>>>(args);
        break;
    }
}

template <template <bool> class Access>
static void launch_fused_aligned_as(const StepArgs &args, int rows, bool nontemporal, hipStream_t stream) {
    const uint32_t blocks = (args.tiles_total + 3) / 4;
#define DPA_ROWS(R)                                                                   \
    case R:                                                                           \
        if (nontemporal)                                                              \
            k_fused_aligned<R, Access<true>><<<blocks, 256, 0, stream>>>(args);       \
        else                                                                          \
            k_fused_aligned<R, Access<false>><<<blocks, 256, 0, stream>>>(args);      \
        break;
    switch (rows) {
        DPA_ROWS(4)
        DPA_ROWS(6)
        DPA_ROWS(8)
        DPA_ROWS(10)
        DPA_ROWS(12)
    default:
        k_fused_aligned<4, Access<true>><<<blocks, 256, 0, stream>>>(args);
        break;
    }
#undef DPA_ROWS
}

template <int R, typename Access>
static void launch_batch_as(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                            uint32_t n_tiles, const InjDesc *injections, bool direct, uint32_t blocks,
                            hipStream_t stream) {
    if (direct)
        k_fused_batch_direct<R, Access><<<(n_tiles + 3) / 4, 256, 0, stream>>>(instances, slots, owners, n_tiles,
                                                                               injections);
    else
        k_fused_batch<R, Access><<<blocks, 256, 0, stream>>>(instances, slots, owners, n_tiles, injections);
}

template <int R>
static void launch_batch_rows(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                              uint32_t n_tiles, const InjDesc *injections, bool nontemporal, bool buffer, bool direct,
                              uint32_t blocks, hipStream_t stream) {
    if (buffer) {
        if (nontemporal)
            launch_batch_as<R, BufferAccess<true>>(instances, slots, owners, n_tiles, injections, direct, blocks, stream);
        else
            launch_batch_as<R, BufferAccess<false>>(instances, slots, owners, n_tiles, injections, direct, blocks,
                                                    stream);
    } else if (nontemporal) {
        launch_batch_as<R, Pointer32<true>>(instances, slots, owners, n_tiles, injections, direct, blocks, stream);
    } else {
        launch_batch_as<R, Pointer32<false>>(instances, slots, owners, n_tiles, injections, direct, blocks, stream);
    }
}

void launch_fused_batch(const InstanceDesc *instances, const StepSlot *slots, const uint32_t *owners,
                        uint32_t n_tiles, const InjDesc *injections, int rows, bool nontemporal, bool buffer,
                        bool direct, uint32_t blocks, hipStream_t stream) {
    switch (rows) {
    case 6:
        launch_batch_rows<6>(instances, slots, owners, n_tiles, injections, nontemporal, buffer, direct, blocks,
                             stream);
        break;
    case 8:
        launch_batch_rows<8>(instances, slots, owners, n_tiles, injections, nontemporal, buffer, direct, blocks,
                             stream);
        break;
    default:
        launch_batch_rows<4>(instances, slots, owners, n_tiles, injections, nontemporal, buffer, direct, blocks,
                             stream);
        break;
    }
}

uint32_t batch_resident_blocks(int rows, bool nontemporal) {
    int device = 0, per_cu = 0;
    (void)hipGetDevice(&device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 1024;
    const void *kernel = nullptr;
    switch (rows) {
    case 6:
        kernel = nontemporal ? (const void *)k_fused_batch<6, Pointer32<true>> : (const void *)k_fused_batch<6, Pointer32<false>>;
        break;
    case 8:
        kernel = nontemporal ? (const void *)k_fused_batch<8, Pointer32<true>> : (const void *)k_fused_batch<8, Pointer32<false>>;
        break;
    default:
        kernel = nontemporal ? (const void *)k_fused_batch<4, Pointer32<true>> : (const void *)k_fused_batch<4, Pointer32<false>>;
        break;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu <= 0)
        per_cu = 4;
    return uint32_t(per_cu) * uint32_t(prop.multiProcessorCount);
}

bool buffer_addressing_fits(uint64_t cells, uint32_t nx, uint32_t ny) {
    // largest offset from a plane's base: 2 components + the next-plane step
    // in the scalar offset, plus a lane's in-plane offset
    const uint64_t comp = cells * sizeof(float), plane = uint64_t(nx) * ny * sizeof(float);
    return 2 * comp + 2 * plane < 0xFFFFFFFFull;
}

void launch_fused_aligned(const StepArgs &args, int rows, bool nontemporal, bool buffer, hipStream_t stream) {
    if (buffer && buffer_addressing_fits(args.cells, args.nx, args.ny))
        launch_fused_aligned_as<BufferAccess>(args, rows, nontemporal, stream);
    // 32-bit byte offsets while one component fits in 4 GiB (1 Gi cells).
    else if (args.cells * sizeof(float) <= 0xFFFFFFFFull)
        launch_fused_aligned_as<Pointer32>(args, rows, nontemporal, stream);
    else
        launch_fused_aligned_as<Pointer64>(args, rows, nontemporal, stream);
}

void launch_update_h(const StepArgs &args, int direction, hipStream_t stream) {
    dim3 grid((args.nx + 63) / 64, (args.ny + 3) / 4, args.nz), block(64, 4, 1);
    if (direction > 0)
        k_update_h<1><<<grid, block, 0, stream>>>(args);
    else
        k_update_h<-1><<<grid, block, 0, stream>>>(args);
}

void launch_update_e(const StepArgs &args, int direction, hipStream_t stream) {
    dim3 grid((args.nx + 63) / 64, (args.ny + 3) / 4, args.nz), block(64, 4, 1);
    if (direction > 0)
        k_update_e<1><<<grid, block, 0, stream>>>(args);
    else
        k_update_e<-1><<<grid, block, 0, stream>>>(args);
}

void launch_inject(float *field, uint64_t cells, uint32_t nx, uint32_t ny, const InjDesc &desc,
                   const InjDesc *, const float *weights, hipStream_t stream) {
    dim3 grid((desc.extent[0] + 255) / 256, desc.extent[1], desc.extent[2]);
    k_inject<<<grid, 256, 0, stream>>>(field + uint64_t(desc.component) * cells, nx, ny, desc, weights);
}

void launch_intensity_f32(const double *sum, float *out, uint64_t cells, hipStream_t stream) {
    const uint32_t blocks = uint32_t((cells + 255) / 256);
    k_intensity_f32<<<blocks, 256, 0, stream>>>(sum, out, cells);
}

void launch_intensity(const float *e, double *intensity, uint64_t cells, hipStream_t stream) {
    const uint32_t blocks = uint32_t((cells + 255) / 256);
    k_intensity<<<blocks, 256, 0, stream>>>(e, intensity, cells);
}

uint32_t energy_blocks(uint32_t nx, uint32_t ny, uint32_t nz) {
    const uint64_t n = uint64_t(nx) * ny * nz;
    const uint64_t wanted = (n + kEnergyThreads - 1) / kEnergyThreads;
    return uint32_t(wanted < 4096 ? wanted : 4096);
}

uint32_t peak_blocks(uint64_t cells) {
    const uint64_t wanted = (cells + kPeakThreads * 16 - 1) / (kPeakThreads * 16);
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(wanted, 4096)));
}

uint32_t launch_measure_peak(const float *e, const float *h, const double *intensity, uint64_t steps, uint64_t cells,
                             float *partials, hipStream_t stream) {
    const uint32_t blocks = peak_blocks(cells);
    k_measure_peak<<<blocks, kPeakThreads, 0, stream>>>(e, h, intensity, float(steps ? steps : 1), cells, partials);
    return blocks;
}

uint32_t launch_energy(const float *e, const float *h, const uint8_t *material,
                       const float2 *eps_mu, const float *widths, uint32_t nx, uint32_t ny,
                       uint32_t nz, double *partials, hipStream_t stream) {
    const uint32_t blocks = energy_blocks(nx, ny, nz);
    k_energy<<<blocks, kEnergyThreads, 0, stream>>>(e, h, material, eps_mu, widths, nx, ny, nz,
                                                    partials);
    return blocks;
}

}  // namespace dpa
