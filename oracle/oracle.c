/*
 * CPU oracle: plain-C restatement of kvark/diaphane's src/cpu.rs and of the
 * scene layer it is built on. TEST INFRASTRUCTURE ONLY — see oracle.h.
 *
 * Arithmetic contract: every expression keeps the reference's f32 operation
 * order (Rust evaluates left to right and never contracts a*b+c into an
 * FMA; this file is compiled with -ffp-contract=off). Transcendentals are
 * glibc's expf/sinf/cosf/asinf/logf/powf, which is what Rust's f32 methods
 * call on Linux. Corners are summed in f64 exactly as src/grid.rs:174-189.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/grid.rs:42-48 */
#define C0 299792458.0f
#define MU0 1.2566371e-6f
#define EPS0 8.854188e-12f
#define COURANT_LIMIT 0.57735026f
#define PI_F 3.14159265358979323846f

static __thread char g_error[512];

const char *or_last_error(void) { return g_error; }

static int fail(const char *fmt, ...) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, args);
    va_end(args);
    return -1;
}

/* ---------------------------------------------------------------------- */
/* Spacing (src/grid.rs:150-286)                                          */
/* ---------------------------------------------------------------------- */
typedef struct {
    uint32_t count;
    float *primary; /* count   */
    float *dual;    /* count   */
    float *corners; /* count+1 */
} spacing_t;

static void spacing_free(spacing_t *s) {
    free(s->primary);
    free(s->dual);
    free(s->corners);
    memset(s, 0, sizeof *s);
}

/* Spacing::assemble, src/grid.rs:190-210: dual[i] = ½(w[i−1] + w[i]) with
 * the first cell mirrored. */
static void spacing_dual(spacing_t *s) {
    s->dual = malloc(sizeof(float) * s->count);
    for (uint32_t i = 0; i < s->count; ++i) {
        uint32_t prev = i == 0 ? 0 : i - 1;
        s->dual[i] = 0.5f * (s->primary[prev] + s->primary[i]);
    }
}

/* Spacing::uniform, src/grid.rs:164-175: corners multiplied in f64. */
static void spacing_uniform(spacing_t *s, uint32_t count, float width) {
    s->count = count;
    s->primary = malloc(sizeof(float) * count);
    s->corners = malloc(sizeof(float) * (count + 1));
    for (uint32_t i = 0; i < count; ++i) s->primary[i] = width;
    for (uint32_t i = 0; i <= count; ++i)
        s->corners[i] = (float)((double)i * (double)width);
    spacing_dual(s);
}

/* Spacing::from_widths, src/grid.rs:178-188: corners summed in f64. */
static void spacing_from_widths(spacing_t *s, float *widths, uint32_t count) {
    s->count = count;
    s->primary = widths;
    s->corners = malloc(sizeof(float) * (count + 1));
    double offset = 0.0;
    s->corners[0] = 0.0f;
    for (uint32_t i = 0; i < count; ++i) {
        offset += (double)widths[i];
        s->corners[i + 1] = (float)offset;
    }
    spacing_dual(s);
}

static float spacing_length(const spacing_t *s) { return s->corners[s->count]; }

static float spacing_finest(const spacing_t *s) {
    float v = INFINITY;
    for (uint32_t i = 0; i < s->count; ++i) v = fminf(v, s->primary[i]);
    return v;
}

static float spacing_coarsest(const spacing_t *s) {
    float v = -INFINITY;
    for (uint32_t i = 0; i < s->count; ++i) v = fmaxf(v, s->primary[i]);
    return v;
}

static float spacing_worst_ratio(const spacing_t *s) {
    float worst = 1.0f;
    for (uint32_t i = 0; i + 1 < s->count; ++i) {
        float a = s->primary[i], b = s->primary[i + 1];
        worst = fmaxf(worst, fmaxf(b / a, a / b));
    }
    return worst;
}

/* Spacing::offset_of, src/grid.rs:280-285. */
static float spacing_offset_of(const spacing_t *s, float coordinate) {
    uint32_t last = s->count - 1;
    float fl = fmaxf(floorf(coordinate), 0.0f);
    /* `as usize` saturates; clamp before casting so the cast is defined */
    uint64_t index = fl >= (float)last ? last : (uint64_t)fl;
    return s->corners[index] + (coordinate - (float)index) * s->primary[index];
}

/* Spacing::position, src/grid.rs:253-255. */
static float spacing_position(const spacing_t *s, float coordinate) {
    return spacing_offset_of(s, coordinate) - 0.5f * spacing_length(s);
}

/* Spacing::coordinate, src/grid.rs:260-272: partition_point over corners. */
static float spacing_coordinate(const spacing_t *s, float position) {
    float offset = position + 0.5f * spacing_length(s);
    /* first index whose corner is > offset */
    uint32_t lo = 0, hi = s->count + 1;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (s->corners[mid] <= offset)
            lo = mid + 1;
        else
            hi = mid;
    }
    uint32_t index = lo == 0 ? 0 : lo - 1;
    if (index > s->count - 1) index = s->count - 1;
    return (float)index + (offset - s->corners[index]) / s->primary[index];
}

/* grade(), src/grid.rs:791-888. */
static int grade(spacing_t *out, int axis, float length, float base,
                 const dp_refinement *all, uint32_t all_count, float max_ratio) {
    const dp_refinement **refs = malloc(sizeof(*refs) * (all_count + 1));
    uint32_t n = 0;
    for (uint32_t i = 0; i < all_count; ++i)
        if (all[i].size[axis] > 0.0f) refs[n++] = &all[i];
    if (n == 0) {
        free(refs);
        float cells = fmaxf(roundf(length / base), 2.0f);
        spacing_uniform(out, (uint32_t)cells, base);
        return 0;
    }
    float finest = base;
    for (uint32_t i = 0; i < n; ++i) finest = fminf(finest, refs[i]->cell_size);
    if (!(finest > 0.0f && isfinite(finest))) {
        free(refs);
        return fail("refinement cell sizes must be positive and finite, got %g", finest);
    }
    double raw = ceil((double)(length / finest * 4.0f));
    size_t samples = raw < 64.0 ? 64 : raw > (double)(1u << 20) ? (1u << 20) : (size_t)raw;
    float step = length / (float)samples;

    size_t tlen = samples + 1;
    float *target = malloc(sizeof(float) * tlen);
    for (size_t i = 0; i < tlen; ++i) {
        float width = base;
        float position = (float)i * step - 0.5f * length;
        for (uint32_t r = 0; r < n; ++r) {
            float half = 0.5f * refs[r]->size[axis];
            if (fabsf(position - refs[r]->center[axis]) <= half)
                width = fminf(width, refs[r]->cell_size);
        }
        target[i] = width;
    }
    free(refs);
    float slope = (max_ratio - 1.0f) * step;
    for (size_t i = 1; i < tlen; ++i) target[i] = fminf(target[i], target[i - 1] + slope);
    for (size_t i = tlen - 1; i-- > 0;) target[i] = fminf(target[i], target[i + 1] + slope);

    size_t cap = 1024, len = 0;
    float *widths = malloc(sizeof(float) * cap);
    float offset = 0.0f;
    while (offset < length) {
        size_t index = (size_t)(offset / step);
        if (index > tlen - 1) index = tlen - 1;
        if (len == cap) widths = realloc(widths, sizeof(float) * (cap *= 2));
        widths[len++] = target[index];
        offset += target[index];
        if (len > (1u << 16)) {
            free(widths);
            free(target);
            return fail("grading the %d axis ran to %zu cells; the refinements are "
                        "finer than the domain can carry", axis, len);
        }
    }
    free(target);
    for (size_t i = 1; i < len; ++i) widths[i] = fminf(widths[i], widths[i - 1] * max_ratio);
    for (size_t i = len - 1; i-- > 0;) widths[i] = fminf(widths[i], widths[i + 1] * max_ratio);

    float total = -0.0f;
    for (size_t i = 0; i < len; ++i) total += widths[i];
    if (len > 2) {
        float without = total - widths[len - 1];
        if (fabsf(length - without) < fabsf(total - length)) {
            len -= 1;
            total = without;
        }
    }
    float scale = length / total;
    for (size_t i = 0; i < len; ++i) widths[i] *= scale;
    if (len < 2) {
        free(widths);
        return fail("an axis needs at least 2 cells, got %zu", len);
    }
    spacing_from_widths(out, widths, (uint32_t)len);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Grid (src/grid.rs:345-773)                                             */
/* ---------------------------------------------------------------------- */
typedef struct {
    uint32_t extent[3];
    float courant;
    spacing_t spacing[3];
} grid_t;

static void grid_free(grid_t *g) {
    for (int a = 0; a < 3; ++a) spacing_free(&g->spacing[a]);
}

/* Grid::build (src/grid.rs:466-500) followed by Grid::validate (:747-773),
 * which both solvers call first (src/cpu.rs:101, src/gpu.rs:191). */
static int grid_build(grid_t *g, const dp_grid_spec *spec) {
    memset(g, 0, sizeof *g);
    if (!(spec->cell_size > 0.0f && isfinite(spec->cell_size)))
        return fail("cell size must be positive and finite, got %g", spec->cell_size);
    if (!(spec->max_ratio > 1.0f && isfinite(spec->max_ratio)))
        return fail("the growth cap must exceed 1, got %g", spec->max_ratio);
    for (int a = 0; a < 3; ++a) {
        if (spec->refinement_count == 0) {
            if (spec->extent[a] < 2) {
                grid_free(g);
                return fail("an axis needs at least 2 cells, got %u", spec->extent[a]);
            }
            spacing_uniform(&g->spacing[a], spec->extent[a], spec->cell_size);
        } else {
            float length = (float)spec->extent[a] * spec->cell_size;
            if (grade(&g->spacing[a], a, length, spec->cell_size, spec->refinements,
                      spec->refinement_count, spec->max_ratio) != 0) {
                grid_free(g);
                return -1;
            }
        }
        g->extent[a] = g->spacing[a].count;
    }
    g->courant = spec->courant;
    if (!(g->courant > 0.0f && g->courant <= COURANT_LIMIT)) {
        grid_free(g);
        return fail("Courant number %g is outside (0, %g]; the scheme would diverge",
                    spec->courant, COURANT_LIMIT);
    }
    return 0;
}

/* Grid::reference_cell_size, src/grid.rs:575-586. */
static float grid_reference_cell_size(const grid_t *g) {
    float sum = -0.0f;
    for (int a = 0; a < 3; ++a) {
        float f = spacing_finest(&g->spacing[a]);
        sum += 1.0f / (f * f);
    }
    return sqrtf(3.0f / sum);
}

/* Grid::time_step, src/grid.rs:597-599. */
static float grid_time_step(const grid_t *g) {
    return g->courant * grid_reference_cell_size(g) / C0;
}

/* Grid::gains, src/grid.rs:616-619. */
static void grid_gains(const grid_t *g, const float *widths, uint32_t n, float *out) {
    float travel = C0 * grid_time_step(g);
    for (uint32_t i = 0; i < n; ++i) out[i] = travel / widths[i];
}

static float grid_coarsest(const grid_t *g) {
    float v = -INFINITY;
    for (int a = 0; a < 3; ++a) v = fmaxf(v, spacing_coarsest(&g->spacing[a]));
    return v;
}

/* Grid::cell_containing, src/grid.rs:710-714. */
static void grid_cell_containing(const grid_t *g, const float p[3], uint32_t out[3]) {
    for (int a = 0; a < 3; ++a) {
        float cell = spacing_coordinate(&g->spacing[a], p[a]);
        float fl = fmaxf(floorf(cell), 0.0f); /* fmaxf drops a NaN like f32::max */
        uint32_t last = g->extent[a] - 1;
        out[a] = fl >= (float)last ? last : (uint32_t)fl;
    }
}

/* Grid::cell_center, src/grid.rs:718-720. */
static void grid_cell_center(const grid_t *g, const uint32_t c[3], float out[3]) {
    for (int a = 0; a < 3; ++a)
        out[a] = spacing_position(&g->spacing[a], (float)c[a] + 0.5f);
}

/* Grid::contains, src/grid.rs:723-726. */
static int grid_contains(const grid_t *g, const float p[3]) {
    for (int a = 0; a < 3; ++a)
        if (!(fabsf(p[a]) <= 0.5f * spacing_length(&g->spacing[a]))) return 0;
    return 1;
}

int or_grid_resolve(const dp_grid_spec *spec, uint32_t extent_out[3], float *widths_out,
                    size_t capacity, float *time_step_out) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return -1;
    memcpy(extent_out, g.extent, sizeof g.extent);
    if (widths_out) {
        size_t need = (size_t)g.extent[0] + g.extent[1] + g.extent[2];
        if (capacity < need) {
            grid_free(&g);
            return fail("widths buffer holds %zu, need %zu", capacity, need);
        }
        size_t at = 0;
        for (int a = 0; a < 3; ++a)
            for (uint32_t i = 0; i < g.extent[a]; ++i) widths_out[at++] = g.spacing[a].primary[i];
    }
    if (time_step_out) *time_step_out = grid_time_step(&g);
    grid_free(&g);
    return 0;
}

int or_grid_dual(const dp_grid_spec *spec, float *out, size_t capacity) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return -1;
    size_t need = (size_t)g.extent[0] + g.extent[1] + g.extent[2], at = 0;
    if (capacity < need) {
        grid_free(&g);
        return fail("buffer holds %zu, need %zu", capacity, need);
    }
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < g.extent[a]; ++i) out[at++] = g.spacing[a].dual[i];
    grid_free(&g);
    return 0;
}

/* Grid::packed_geometry, src/grid.rs:666-678. */
int or_grid_packed_geometry(const dp_grid_spec *spec, float *out, size_t capacity) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return -1;
    size_t span = (size_t)g.extent[0] + g.extent[1] + g.extent[2];
    if (capacity < 3 * span) {
        grid_free(&g);
        return fail("buffer holds %zu, need %zu", capacity, 3 * span);
    }
    size_t at = 0;
    for (int a = 0; a < 3; ++a) {
        grid_gains(&g, g.spacing[a].primary, g.extent[a], out + at);
        at += g.extent[a];
    }
    for (int a = 0; a < 3; ++a) {
        grid_gains(&g, g.spacing[a].dual, g.extent[a], out + at);
        at += g.extent[a];
    }
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < g.extent[a]; ++i)
            out[at++] = spacing_position(&g.spacing[a], (float)i + 0.5f);
    grid_free(&g);
    return 0;
}

float or_grid_coordinate(const dp_grid_spec *spec, int axis, float position) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return NAN;
    float v = spacing_coordinate(&g.spacing[axis], position);
    grid_free(&g);
    return v;
}

float or_grid_position(const dp_grid_spec *spec, int axis, float coordinate) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return NAN;
    float v = spacing_position(&g.spacing[axis], coordinate);
    grid_free(&g);
    return v;
}

/* Grid::cell_lookup, src/grid.rs:635-646: per axis, `samples` fractional
 * cell coordinates of positions from wall to wall, f32 throughout. */
int or_grid_cell_lookup(const dp_grid_spec *spec, uint32_t samples, float *out, size_t capacity) {
    if (samples < 2) return fail("a lookup needs at least two samples");
    if (capacity < (size_t)3 * samples) return fail("lookup buffer too small");
    grid_t g;
    if (grid_build(&g, spec) != 0) return -1;
    size_t at = 0;
    for (int a = 0; a < 3; ++a) {
        float length = spacing_length(&g.spacing[a]);
        for (uint32_t i = 0; i < samples; ++i) {
            float fraction = (float)i / (float)(samples - 1);
            out[at++] = spacing_coordinate(&g.spacing[a], (fraction - 0.5f) * length);
        }
    }
    grid_free(&g);
    return 0;
}

int or_grid_cell_containing(const dp_grid_spec *spec, const float position[3], uint32_t out[3]) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return -1;
    grid_cell_containing(&g, position, out);
    grid_free(&g);
    return 0;
}

/* numerical_phase_velocity, src/grid.rs:966-1012 (f64 bisection). */
float or_numerical_phase_velocity(float courant_f, float cells_f, const float direction[3]) {
    double courant = courant_f, cells = cells_f;
    double norm = 0.0;
    for (int a = 0; a < 3; ++a) norm += (double)direction[a] * (double)direction[a];
    norm = sqrt(norm);
    double unit[3];
    for (int a = 0; a < 3; ++a) unit[a] = (double)direction[a] / norm;
    double s = sin(M_PI * courant / cells) / courant;
    double target = s * s;
    double ideal = 2.0 * M_PI / cells;
    double lo = 0.5 * ideal, hi = 2.0 * ideal;
    for (int it = 0; it < 80; ++it) {
        double mid = 0.5 * (lo + hi);
        double r = 0.0;
        for (int a = 0; a < 3; ++a) {
            double v = sin(0.5 * mid * unit[a]);
            r += v * v;
        }
        if (r - target < 0.0)
            lo = mid;
        else
            hi = mid;
    }
    return (float)(ideal / (0.5 * (lo + hi)));
}

/* ---------------------------------------------------------------------- */
/* Materials (src/material.rs:100-122)                                    */
/* ---------------------------------------------------------------------- */
typedef struct {
    float electric_gain, electric_loss, magnetic_gain, magnetic_loss;
} coeff_t;

static coeff_t material_coefficients(const dp_material *m, float time_step) {
    coeff_t c;
    if (isinf(m->conductivity)) {
        c.electric_gain = 0.0f;
        c.electric_loss = 1.0f;
        c.magnetic_gain = 1.0f / m->relative_permeability;
        c.magnetic_loss = 0.0f;
        return c;
    }
    float half_step = 0.5f * time_step;
    c.electric_gain = 1.0f / m->relative_permittivity;
    c.electric_loss = half_step * m->conductivity / (EPS0 * m->relative_permittivity);
    c.magnetic_gain = 1.0f / m->relative_permeability;
    c.magnetic_loss = half_step * m->magnetic_loss / (MU0 * m->relative_permeability);
    return c;
}

int or_material_coefficients(const dp_material *materials, uint32_t count, float time_step,
                             float *out) {
    for (uint32_t i = 0; i < count; ++i) {
        coeff_t c = material_coefficients(&materials[i], time_step);
        out[4 * i + 0] = c.electric_gain;
        out[4 * i + 1] = c.electric_loss;
        out[4 * i + 2] = c.magnetic_gain;
        out[4 * i + 3] = c.magnetic_loss;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Absorber (src/boundary.rs:100-145)                                     */
/* ---------------------------------------------------------------------- */
typedef struct {
    float *integer[3];
    float *half[3];
    float peak;
} absorber_t;

static void absorber_free(absorber_t *p) {
    for (int a = 0; a < 3; ++a) {
        free(p->integer[a]);
        free(p->half[a]);
    }
}

static float absorber_profile(float position, uint32_t n, uint32_t thickness, float peak) {
    float outer = (float)(n - 1);
    float t = (float)thickness;
    float depth_low = (t - position) / t;
    float depth_high = (t - (outer - position)) / t;
    float depth = fmaxf(depth_low, depth_high);
    depth = depth < 0.0f ? 0.0f : depth > 1.0f ? 1.0f : depth;
    return peak * powf(depth, 3.0f);
}

static int absorber_build(absorber_t *p, const grid_t *g, const dp_scene *scene) {
    memset(p, 0, sizeof *p);
    uint32_t thickness = 0;
    float target = 1.0f;
    if (scene->boundary == DP_BOUNDARY_ABSORBING) {
        thickness = scene->absorber_thickness;
        target = scene->target_reflection;
    } else if (scene->boundary != DP_BOUNDARY_PEC) {
        return fail("unknown boundary kind %u", scene->boundary);
    }
    for (int a = 0; a < 3; ++a) {
        p->integer[a] = calloc(g->extent[a], sizeof(float));
        p->half[a] = calloc(g->extent[a], sizeof(float));
    }
    if (thickness == 0) return 0;
    for (int a = 0; a < 3; ++a)
        if (!(2 * (uint64_t)thickness < g->extent[a])) {
            absorber_free(p);
            return fail("absorbing layers of %u cells do not fit in axis %d of %u cells",
                        thickness, a, g->extent[a]);
        }
    if (!(target > 0.0f && target < 1.0f)) {
        absorber_free(p);
        return fail("target reflection %g must lie in (0, 1)", target);
    }
    const float order = 3.0f;
    float peak = -(order + 1.0f) * logf(target) * g->courant / (4.0f * (float)thickness);
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < g->extent[a]; ++i) {
            p->integer[a][i] = absorber_profile((float)i, g->extent[a], thickness, peak);
            p->half[a][i] = absorber_profile((float)i + 0.5f, g->extent[a], thickness, peak);
        }
    float best = 0.0f;
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < g->extent[a]; ++i)
            best = fmaxf(best, fmaxf(p->integer[a][i], p->half[a][i]));
    p->peak = best;
    return 0;
}

int or_absorber_packed(const dp_scene *scene, float *out, size_t capacity) {
    grid_t g;
    if (grid_build(&g, &scene->grid) != 0) return -1;
    absorber_t p;
    if (absorber_build(&p, &g, scene) != 0) {
        grid_free(&g);
        return -1;
    }
    size_t span = (size_t)g.extent[0] + g.extent[1] + g.extent[2];
    if (capacity < 2 * span) {
        absorber_free(&p);
        grid_free(&g);
        return fail("buffer holds %zu, need %zu", capacity, 2 * span);
    }
    size_t at = 0;
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < g.extent[a]; ++i) out[at++] = p.integer[a][i];
    for (int a = 0; a < 3; ++a)
        for (uint32_t i = 0; i < g.extent[a]; ++i) out[at++] = p.half[a][i];
    absorber_free(&p);
    grid_free(&g);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Waveforms and injections (src/source.rs)                               */
/* ---------------------------------------------------------------------- */
float or_waveform_evaluate(const dp_waveform *w, float time) {
    switch (w->kind) {
    case DP_WAVE_RICKER: {
        float arg = PI_F * w->frequency * (time - w->delay);
        float arg2 = arg * arg;
        return (1.0f - 2.0f * arg2) * expf(-arg2);
    }
    case DP_WAVE_GAUSSIAN_PULSE: {
        const float fwhm_to_sigma = 0.4246609f;
        float sigma = fwhm_to_sigma * w->cycles / w->frequency;
        float tau = time - w->delay;
        float envelope = expf(-0.5f * (tau / sigma) * (tau / sigma));
        return envelope * sinf(2.0f * PI_F * w->frequency * tau);
    }
    case DP_WAVE_CONTINUOUS_WAVE: {
        if (time <= 0.0f) return 0.0f;
        float ramp = w->cycles / w->frequency;
        float envelope = time >= ramp ? 1.0f : 0.5f * (1.0f - cosf(PI_F * time / ramp));
        return envelope * sinf(2.0f * PI_F * w->frequency * time);
    }
    case DP_WAVE_DIFF_GAUSSIAN: {
        float u = (time - w->delay) / w->width;
        return -u * expf(0.5f * (1.0f - u * u));
    }
    default:
        return 0.0f;
    }
}

/* Waveform::dominant_frequency, src/source.rs:118-126. */
static float waveform_dominant_frequency(const dp_waveform *w) {
    if (w->kind == DP_WAVE_DIFF_GAUSSIAN) return 1.0f / (2.0f * PI_F * w->width);
    return w->frequency;
}

static dp_injection silent(dp_injection inj) {
    inj.value = 0.0f;
    return inj;
}

/* Source::injections, src/source.rs:280-403. */
static void source_injections(const dp_source *s, const grid_t *g, float time, dp_injection out[2]) {
    float value = s->amplitude * or_waveform_evaluate(&s->waveform, time);
    dp_injection inj;
    memset(&inj, 0, sizeof inj);
    if (s->shape == DP_SOURCE_POINT) {
        float p[3];
        for (int a = 0; a < 3; ++a) p[a] = s->at[a] + s->velocity[a] * time;
        grid_cell_containing(g, p, inj.origin);
        inj.extent[0] = inj.extent[1] = inj.extent[2] = 1;
        inj.component = s->polarization;
        inj.magnetic = 0;
        inj.value = grid_contains(g, p) ? value : 0.0f;
        out[0] = inj;
        out[1] = silent(inj);
    } else if (s->shape == DP_SOURCE_SHEET) {
        int a = (int)s->axis;
        float p[3] = {0.0f, 0.0f, 0.0f};
        p[a] = s->offset;
        uint32_t cell[3];
        grid_cell_containing(g, p, cell);
        for (int t = 0; t < 3; ++t) inj.extent[t] = g->extent[t];
        inj.origin[a] = cell[a];
        inj.extent[a] = 1;
        inj.inverse_waist_squared = 1.0f / (s->waist * s->waist);
        inj.component = s->polarization;
        inj.value = value;
        out[0] = inj;
        out[1] = silent(inj);
    } else {
        int travel = (int)s->axis;
        float p[3] = {0.0f, 0.0f, 0.0f};
        p[travel] = s->offset;
        uint32_t cell[3];
        grid_cell_containing(g, p, cell);
        uint32_t plane = cell[travel] > 1 ? cell[travel] : 1;
        float spacing = spacing_finest(&g->spacing[travel]);
        float time_step = grid_time_step(g);
        float gain = C0 * time_step / spacing;
        float omega = 2.0f * PI_F * waveform_dominant_frequency(&s->waveform);
        float half_k_spacing = asinf(sinf(0.5f * omega * time_step) / gain);
        float retard = half_k_spacing / omega;
        int pol = (int)s->polarization;
        int h_axis = 3 - travel - pol;
        float h_sign = ((pol + 1) % 3 == travel) ? -1.0f : 1.0f;
        dp_injection e = inj, h = inj;
        for (int t = 0; t < 3; ++t) e.extent[t] = h.extent[t] = g->extent[t];
        e.extent[travel] = h.extent[travel] = 1;
        e.origin[travel] = plane;
        h.origin[travel] = plane - 1;
        e.component = (uint32_t)pol;
        e.magnetic = 0;
        e.value = gain * (s->amplitude * or_waveform_evaluate(&s->waveform,
                                                             time - 0.5f * time_step + retard));
        h.component = (uint32_t)h_axis;
        h.magnetic = 1;
        h.value = h_sign * gain * (s->amplitude * or_waveform_evaluate(&s->waveform, time - time_step));
        out[0] = e;
        out[1] = h;
    }
}

/* Injection::weight, src/source.rs:441-455. */
static float injection_weight(const dp_injection *inj, const float position[3]) {
    if (inj->inverse_waist_squared == 0.0f) return 1.0f;
    float r2 = 0.0f;
    for (int a = 0; a < 3; ++a)
        if (inj->extent[a] > 1) {
            float off = position[a] - inj->center[a];
            r2 += off * off;
        }
    return expf(-r2 * inj->inverse_waist_squared);
}

int or_scene_injections(const dp_scene *scene, float time, dp_injection *out, size_t capacity) {
    if (capacity < 2 * (size_t)scene->source_count) return fail("injection buffer too small");
    grid_t g;
    if (grid_build(&g, &scene->grid) != 0) return -1;
    for (uint32_t s = 0; s < scene->source_count; ++s)
        source_injections(&scene->sources[s], &g, time, out + 2 * s);
    grid_free(&g);
    return 0;
}

float or_injection_weight(const dp_scene *scene, const dp_injection *inj, const uint32_t coord[3]) {
    grid_t g;
    if (grid_build(&g, &scene->grid) != 0) return NAN;
    float p[3];
    grid_cell_center(&g, coord, p);
    float w = injection_weight(inj, p);
    grid_free(&g);
    return w;
}

/* ---------------------------------------------------------------------- */
/* Scene rasterization (src/scene.rs:60-81, 181-206)                      */
/* ---------------------------------------------------------------------- */
static int shape_contains(const dp_shape *s, const float p[3]) {
    switch (s->kind) {
    case DP_SHAPE_BLOCK:
        for (int a = 0; a < 3; ++a)
            if (!(fabsf(p[a] - s->center[a]) <= 0.5f * s->size[a])) return 0;
        return 1;
    case DP_SHAPE_SPHERE: {
        float d2 = -0.0f;
        for (int a = 0; a < 3; ++a) {
            float off = p[a] - s->center[a];
            d2 += off * off;
        }
        return d2 <= s->radius * s->radius;
    }
    case DP_SHAPE_SLAB:
        return fabsf(p[s->axis] - s->offset) <= 0.5f * s->thickness;
    default:
        return 0;
    }
}

static void rasterize(const dp_scene *scene, const grid_t *g, uint32_t *indices) {
    uint64_t n = (uint64_t)g->extent[0] * g->extent[1] * g->extent[2];
    memset(indices, 0, sizeof(uint32_t) * n);
    if (scene->shape_count == 0) return;
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < (int64_t)g->extent[2]; ++z)
        for (uint32_t y = 0; y < g->extent[1]; ++y)
            for (uint32_t x = 0; x < g->extent[0]; ++x) {
                uint32_t c[3] = {x, y, (uint32_t)z};
                float p[3];
                grid_cell_center(g, c, p);
                uint64_t i = ((uint64_t)z * g->extent[1] + y) * g->extent[0] + x;
                for (uint32_t s = 0; s < scene->shape_count; ++s)
                    if (shape_contains(&scene->shapes[s], p)) indices[i] = scene->shapes[s].material;
            }
}

int or_scene_material_indices(const dp_scene *scene, uint32_t *out, size_t count) {
    grid_t g;
    if (grid_build(&g, &scene->grid) != 0) return -1;
    uint64_t n = (uint64_t)g.extent[0] * g.extent[1] * g.extent[2];
    if (count != n) {
        grid_free(&g);
        return fail("index buffer holds %zu, grid has %llu cells", count, (unsigned long long)n);
    }
    rasterize(scene, &g, out);
    grid_free(&g);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Simulation (src/cpu.rs)                                                */
/* ---------------------------------------------------------------------- */
struct or_sim {
    grid_t grid;
    uint64_t cells;
    float *electric; /* 3 × cells */
    float *magnetic; /* 3 × cells */
    uint32_t *material_index;
    dp_material *materials;
    coeff_t *coefficients;
    uint32_t material_count;
    float *magnetic_gains[3];
    float *electric_gains[3];
    float *ones;
    absorber_t absorber;
    float *intensity;
    double *intensity64;
    int accumulating;
    uint64_t accumulated;
    dp_source *sources;
    uint32_t source_count;
    uint64_t step;
    int threads;
    /* the f64 track: the same update on double fields (or_sim_f64_*) */
    double *electric64, *magnetic64;
    uint64_t step64;
};

or_sim *or_sim_create(const dp_scene *scene) {
    or_sim *s = calloc(1, sizeof *s);
    if (grid_build(&s->grid, &scene->grid) != 0) {
        free(s);
        return NULL;
    }
    if (scene->material_count == 0) {
        grid_free(&s->grid);
        free(s);
        fail("the material table is empty; slot 0 must be vacuum");
        return NULL;
    }
    for (uint32_t i = 0; i < scene->shape_count; ++i)
        if (scene->shapes[i].material >= scene->material_count) {
            grid_free(&s->grid);
            free(s);
            fail("shape references material %u but the table has %u", scene->shapes[i].material,
                 scene->material_count);
            return NULL;
        }
    if (absorber_build(&s->absorber, &s->grid, scene) != 0) {
        grid_free(&s->grid);
        free(s);
        return NULL;
    }
    const uint32_t *e = s->grid.extent;
    s->cells = (uint64_t)e[0] * e[1] * e[2];
    s->electric = calloc(3 * s->cells, sizeof(float));
    s->magnetic = calloc(3 * s->cells, sizeof(float));
    s->material_index = malloc(sizeof(uint32_t) * s->cells);
    rasterize(scene, &s->grid, s->material_index);
    s->material_count = scene->material_count;
    s->materials = malloc(sizeof(dp_material) * s->material_count);
    memcpy(s->materials, scene->materials, sizeof(dp_material) * s->material_count);
    s->coefficients = malloc(sizeof(coeff_t) * s->material_count);
    float dt = grid_time_step(&s->grid);
    for (uint32_t i = 0; i < s->material_count; ++i)
        s->coefficients[i] = material_coefficients(&s->materials[i], dt);
    for (int a = 0; a < 3; ++a) {
        s->magnetic_gains[a] = malloc(sizeof(float) * e[a]);
        s->electric_gains[a] = malloc(sizeof(float) * e[a]);
        grid_gains(&s->grid, s->grid.spacing[a].primary, e[a], s->magnetic_gains[a]);
        grid_gains(&s->grid, s->grid.spacing[a].dual, e[a], s->electric_gains[a]);
    }
    s->ones = malloc(sizeof(float) * e[0]);
    for (uint32_t i = 0; i < e[0]; ++i) s->ones[i] = 1.0f;
    s->intensity = calloc(s->cells, sizeof(float));
    s->intensity64 = calloc(s->cells, sizeof(double));
    s->source_count = scene->source_count;
    s->sources = malloc(sizeof(dp_source) * (s->source_count + 1));
    if (s->source_count) memcpy(s->sources, scene->sources, sizeof(dp_source) * s->source_count);
    s->threads = 0;
    return s;
}

void or_sim_destroy(or_sim *s) {
    if (!s) return;
    grid_free(&s->grid);
    free(s->electric);
    free(s->magnetic);
    free(s->material_index);
    free(s->materials);
    free(s->coefficients);
    for (int a = 0; a < 3; ++a) {
        free(s->magnetic_gains[a]);
        free(s->electric_gains[a]);
    }
    free(s->ones);
    absorber_free(&s->absorber);
    free(s->intensity);
    free(s->intensity64);
    free(s->electric64);
    free(s->magnetic64);
    free(s->sources);
    free(s);
}

void or_sim_set_threads(or_sim *s, int threads) { s->threads = threads; }
uint64_t or_sim_cells(const or_sim *s) { return s->cells; }
void or_sim_extent(const or_sim *s, uint32_t out[3]) { memcpy(out, s->grid.extent, 12); }
uint64_t or_sim_step_count(const or_sim *s) { return s->step; }
float or_sim_time(const or_sim *s) { return (float)s->step * grid_time_step(&s->grid); }
float *or_sim_electric(or_sim *s) { return s->electric; }
float *or_sim_magnetic(or_sim *s) { return s->magnetic; }
float *or_sim_intensity(or_sim *s) { return s->intensity; }
double *or_sim_intensity_f64(or_sim *s) { return s->intensity64; }
uint64_t or_sim_accumulated_steps(const or_sim *s) { return s->accumulated; }

static int thread_count(const or_sim *s) {
#ifdef _OPENMP
    return s->threads > 0 ? s->threads : omp_get_max_threads();
#else
    (void)s;
    return 1;
#endif
}

/* AbsorbingProfile::magnetic_row / electric_row, src/boundary.rs:173-196:
 * the x profile and the (y, z) constant, summed y-then-z. */
static const float *magnetic_row(const absorber_t *p, int a, uint32_t y, uint32_t z, float *constant) {
    const float *py = (a == 1) ? p->integer[1] : p->half[1];
    const float *pz = (a == 2) ? p->integer[2] : p->half[2];
    *constant = py[y] + pz[z];
    return (a == 0) ? p->integer[0] : p->half[0];
}

static const float *electric_row(const absorber_t *p, int a, uint32_t y, uint32_t z, float *constant) {
    const float *py = (a == 1) ? p->half[1] : p->integer[1];
    const float *pz = (a == 2) ? p->half[2] : p->integer[2];
    *constant = py[y] + pz[z];
    return (a == 0) ? p->half[0] : p->integer[0];
}

/* update_magnetic (src/cpu.rs:283-376) and reverse_magnetic (:489-572).
 * direction +1: H ← ((1−m)H − g·curl)/(1+m); −1: H ← ((1+m)H + g·curl)/(1−m). */
static void sweep_magnetic(or_sim *s, int direction) {
    const uint32_t *extent = s->grid.extent;
    const uint64_t strides[3] = {1, extent[0], (uint64_t)extent[0] * extent[1]};
    int threads = thread_count(s);
    for (int a = 0; a < 3; ++a) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        uint64_t stride_b = strides[b], stride_c = strides[c];
        uint32_t limit[3] = {extent[0], extent[1], extent[2]};
        limit[b] -= 1;
        limit[c] -= 1;
        const float *gain_b = s->magnetic_gains[b], *gain_c = s->magnetic_gains[c];
        const float *along_b = b == 0 ? gain_b : s->ones;
        const float *along_c = c == 0 ? gain_c : s->ones;
        float *target = s->magnetic + (uint64_t)a * s->cells;
        const float *field_b = s->electric + (uint64_t)b * s->cells;
        const float *field_c = s->electric + (uint64_t)c * s->cells;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t zz = 0; zz < (int64_t)limit[2]; ++zz) {
            uint32_t z = (uint32_t)zz;
            for (uint32_t y = 0; y < limit[1]; ++y) {
                uint64_t row = ((uint64_t)z * extent[1] + y) * extent[0];
                float absorb_row;
                const float *absorb_x = magnetic_row(&s->absorber, a, y, z, &absorb_row);
                uint32_t along_row[3] = {0, y, z};
                float row_b = b == 0 ? 1.0f : gain_b[along_row[b]];
                float row_c = c == 0 ? 1.0f : gain_c[along_row[c]];
                for (uint32_t x = 0; x < limit[0]; ++x) {
                    uint64_t i = row + x;
                    float gb = row_b * along_b[x], gc = row_c * along_c[x];
                    float curl = gb * (field_c[i + stride_b] - field_c[i]) -
                                 gc * (field_b[i + stride_c] - field_b[i]);
                    coeff_t k = s->coefficients[s->material_index[i]];
                    float loss = k.magnetic_loss + absorb_row + absorb_x[x];
                    if (direction > 0)
                        target[i] = ((1.0f - loss) * target[i] - k.magnetic_gain * curl) / (1.0f + loss);
                    else
                        target[i] = ((1.0f + loss) * target[i] + k.magnetic_gain * curl) / (1.0f - loss);
                }
            }
        }
        /* the high-wall planes along b and c (src/cpu.rs:335-374) */
        uint32_t last_b = extent[b] - 1, last_c = extent[c] - 1;
        for (int pass = 0; pass < 2; ++pass) {
            int fixed = pass == 0 ? b : c;
            uint32_t at = pass == 0 ? last_b : last_c;
            uint32_t others = pass == 0 ? extent[c] : last_b;
            int other = b + c - fixed;
            for (uint32_t w = 0; w < others; ++w)
                for (uint32_t v = 0; v < extent[a]; ++v) {
                    uint32_t coord[3];
                    coord[fixed] = at;
                    coord[other] = w;
                    coord[a] = v;
                    uint64_t i = ((uint64_t)coord[2] * extent[1] + coord[1]) * extent[0] + coord[0];
                    float next_c = coord[b] + 1 < extent[b] ? field_c[i + stride_b] : 0.0f;
                    float next_b = coord[c] + 1 < extent[c] ? field_b[i + stride_c] : 0.0f;
                    float curl = gain_b[coord[b]] * (next_c - field_c[i]) -
                                 gain_c[coord[c]] * (next_b - field_b[i]);
                    float absorb_row;
                    const float *absorb_x = magnetic_row(&s->absorber, a, coord[1], coord[2], &absorb_row);
                    coeff_t k = s->coefficients[s->material_index[i]];
                    float loss = k.magnetic_loss + absorb_row + absorb_x[coord[0]];
                    if (direction > 0)
                        target[i] = ((1.0f - loss) * target[i] - k.magnetic_gain * curl) / (1.0f + loss);
                    else
                        target[i] = ((1.0f + loss) * target[i] + k.magnetic_gain * curl) / (1.0f - loss);
                }
        }
    }
}

/* update_electric (src/cpu.rs:380-434) and reverse_electric (:441-487). */
static void sweep_electric(or_sim *s, int direction) {
    const uint32_t *extent = s->grid.extent;
    const uint64_t strides[3] = {1, extent[0], (uint64_t)extent[0] * extent[1]};
    int threads = thread_count(s);
    for (int a = 0; a < 3; ++a) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        uint64_t stride_b = strides[b], stride_c = strides[c];
        uint32_t start[3] = {0, 0, 0};
        start[b] = 1;
        start[c] = 1;
        const float *gain_b = s->electric_gains[b], *gain_c = s->electric_gains[c];
        const float *along_b = b == 0 ? gain_b : s->ones;
        const float *along_c = c == 0 ? gain_c : s->ones;
        float *target = s->electric + (uint64_t)a * s->cells;
        const float *field_b = s->magnetic + (uint64_t)b * s->cells;
        const float *field_c = s->magnetic + (uint64_t)c * s->cells;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t zz = start[2]; zz < (int64_t)extent[2]; ++zz) {
            uint32_t z = (uint32_t)zz;
            for (uint32_t y = start[1]; y < extent[1]; ++y) {
                uint64_t row = ((uint64_t)z * extent[1] + y) * extent[0];
                float absorb_row;
                const float *absorb_x = electric_row(&s->absorber, a, y, z, &absorb_row);
                uint32_t along_row[3] = {0, y, z};
                float row_b = b == 0 ? 1.0f : gain_b[along_row[b]];
                float row_c = c == 0 ? 1.0f : gain_c[along_row[c]];
                for (uint32_t x = start[0]; x < extent[0]; ++x) {
                    uint64_t i = row + x;
                    float gb = row_b * along_b[x], gc = row_c * along_c[x];
                    float curl = gb * (field_c[i] - field_c[i - stride_b]) -
                                 gc * (field_b[i] - field_b[i - stride_c]);
                    coeff_t k = s->coefficients[s->material_index[i]];
                    float loss = k.electric_loss + absorb_row + absorb_x[x];
                    if (direction > 0)
                        target[i] = ((1.0f - loss) * target[i] + k.electric_gain * curl) / (1.0f + loss);
                    else
                        target[i] = ((1.0f + loss) * target[i] - k.electric_gain * curl) / (1.0f - loss);
                }
            }
        }
    }
}

/* apply_sources, src/cpu.rs:606-633. */
static void apply_sources(or_sim *s, float time, float sign, int magnetic) {
    const uint32_t *extent = s->grid.extent;
    for (uint32_t k = 0; k < s->source_count; ++k) {
        dp_injection inj[2];
        source_injections(&s->sources[k], &s->grid, time, inj);
        for (int j = 0; j < 2; ++j) {
            const dp_injection *q = &inj[j];
            if (q->value == 0.0f || (int)q->magnetic != magnetic) continue;
            float value = sign * q->value;
            float *target = (q->magnetic ? s->magnetic : s->electric) + (uint64_t)q->component * s->cells;
            for (uint32_t dz = 0; dz < q->extent[2]; ++dz)
                for (uint32_t dy = 0; dy < q->extent[1]; ++dy)
                    for (uint32_t dx = 0; dx < q->extent[0]; ++dx) {
                        uint32_t c[3] = {q->origin[0] + dx, q->origin[1] + dy, q->origin[2] + dz};
                        float p[3];
                        grid_cell_center(&s->grid, c, p);
                        uint64_t i = ((uint64_t)c[2] * extent[1] + c[1]) * extent[0] + c[0];
                        target[i] += value * injection_weight(q, p);
                    }
        }
    }
}

static void clear_intensity(or_sim *s) {
    if (s->accumulated > 0) {
        memset(s->intensity, 0, sizeof(float) * s->cells);
        memset(s->intensity64, 0, sizeof(double) * s->cells);
        s->accumulated = 0;
    }
}

/* advance, src/cpu.rs:174-198. */
static void advance_one(or_sim *s) {
    float time = (float)(s->step + 1) * grid_time_step(&s->grid);
    sweep_magnetic(s, 1);
    apply_sources(s, time, 1.0f, 1);
    sweep_electric(s, 1);
    apply_sources(s, time, 1.0f, 0);
    if (s->accumulating) {
        const float *ex = s->electric, *ey = s->electric + s->cells, *ez = s->electric + 2 * s->cells;
        int threads = thread_count(s);
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t i = 0; i < (int64_t)s->cells; ++i) {
            float sum = 0.0f;
            sum += ex[i] * ex[i];
            sum += ey[i] * ey[i];
            sum += ez[i] * ez[i];
            s->intensity[i] += sum;
            s->intensity64[i] += (double)sum;
        }
        s->accumulated += 1;
    }
    s->step += 1;
}

void or_sim_advance(or_sim *s, uint64_t steps) {
    for (uint64_t i = 0; i < steps; ++i) advance_one(s);
}

/* is_reversible, src/cpu.rs:203-209. */
int or_sim_is_reversible(const or_sim *s) {
    for (uint32_t i = 0; i < s->material_count; ++i)
        if (!(s->coefficients[i].electric_loss == 0.0f && s->coefficients[i].magnetic_loss == 0.0f))
            return 0;
    return s->absorber.peak == 0.0f;
}

/* reverse, src/cpu.rs:253-273. */
int or_sim_reverse(or_sim *s, uint64_t steps) {
    for (uint64_t n = 0; n < steps; ++n) {
        if (!or_sim_is_reversible(s)) {
            fail("this scene is lossy, and running a dissipative update backwards amplifies "
                 "roundoff exponentially; reversal needs Boundary::Pec and no conductive material");
            return DP_ERR_LOSSY;
        }
        if (s->step == 0) {
            fail("cannot step back past the start");
            return DP_ERR_PAST_START;
        }
        clear_intensity(s);
        float time = (float)s->step * grid_time_step(&s->grid);
        apply_sources(s, time, -1.0f, 0);
        sweep_electric(s, -1);
        apply_sources(s, time, -1.0f, 1);
        sweep_magnetic(s, -1);
        s->step -= 1;
    }
    return 0;
}

/* reset, src/cpu.rs:162-171. */
void or_sim_reset(or_sim *s) {
    memset(s->electric, 0, sizeof(float) * 3 * s->cells);
    memset(s->magnetic, 0, sizeof(float) * 3 * s->cells);
    memset(s->intensity, 0, sizeof(float) * s->cells);
    memset(s->intensity64, 0, sizeof(double) * s->cells);
    s->accumulated = 0;
    s->step = 0;
}

/* set_accumulate_intensity, src/cpu.rs:577-582. */
void or_sim_set_accumulate_intensity(or_sim *s, int on) {
    if (on && !s->accumulating) clear_intensity(s);
    s->accumulating = on ? 1 : 0;
}

/* Steppable::restore, src/cpu.rs:736-752. */
int or_sim_restore(or_sim *s, const float *electric, const float *magnetic, size_t count,
                   uint64_t step) {
    clear_intensity(s);
    if (count != 3 * s->cells) {
        fail("snapshot was taken from a different grid");
        return DP_ERR_SIZE;
    }
    memcpy(s->electric, electric, sizeof(float) * count);
    memcpy(s->magnetic, magnetic, sizeof(float) * count);
    s->step = step;
    return 0;
}

/* energy, src/cpu.rs:654-687: sequential f64 sums in z, y, x, axis order. */
void or_sim_energy(const or_sim *s, double *electric, double *magnetic) {
    const uint32_t *extent = s->grid.extent;
    double ue = 0.0, uh = 0.0;
    for (uint32_t z = 0; z < extent[2]; ++z)
        for (uint32_t y = 0; y < extent[1]; ++y) {
            uint64_t row = ((uint64_t)z * extent[1] + y) * extent[0];
            for (uint32_t x = 0; x < extent[0]; ++x) {
                uint64_t i = row + x;
                uint32_t coord[3] = {x, y, z};
                const dp_material *m = &s->materials[s->material_index[i]];
                for (int a = 0; a < 3; ++a) {
                    int b = (a + 1) % 3, c = (a + 2) % 3;
                    float e = s->electric[(uint64_t)a * s->cells + i];
                    float h = s->magnetic[(uint64_t)a * s->cells + i];
                    double ev = (double)s->grid.spacing[a].primary[coord[a]] *
                                (double)s->grid.spacing[b].dual[coord[b]] *
                                (double)s->grid.spacing[c].dual[coord[c]];
                    double hv = (double)s->grid.spacing[a].dual[coord[a]] *
                                (double)s->grid.spacing[b].primary[coord[b]] *
                                (double)s->grid.spacing[c].primary[coord[c]];
                    ue += 0.5 * ev * (double)(m->relative_permittivity * e * e);
                    uh += 0.5 * hv * (double)(m->relative_permeability * h * h);
                }
            }
        }
    *electric = ue;
    *magnetic = uh;
}

/* is_finite, src/cpu.rs:707-713. */
int or_sim_is_finite(const or_sim *s) {
    for (uint64_t i = 0; i < 3 * s->cells; ++i)
        if (!isfinite(s->electric[i]) || !isfinite(s->magnetic[i])) return 0;
    return 1;
}

/* Shared with the tests: frequency_for_resolution (src/grid.rs:742-744). */
float or_grid_frequency_for_resolution(const dp_grid_spec *spec, float cells) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return NAN;
    float f = C0 / (cells * grid_coarsest(&g));
    grid_free(&g);
    return f;
}

/* Spacing::worst_ratio over the grid (src/grid.rs:541-546). */
float or_grid_worst_ratio(const dp_grid_spec *spec) {
    grid_t g;
    if (grid_build(&g, spec) != 0) return NAN;
    float w = 1.0f;
    for (int a = 0; a < 3; ++a) w = fmaxf(w, spacing_worst_ratio(&g.spacing[a]));
    grid_free(&g);
    return w;
}

/* ---------------------------------------------------------------------- */
/* The f64 track: the forward update of src/cpu.rs:174-198 with every      */
/* field value and every arithmetic operation in double. The inputs are    */
/* the same f32 tables (gains, absorber, coefficients, injection values   */
/* and weights), promoted. It measures the f32 rounding error of the      */
/* update itself; it is not bit-compared with anything.                   */
/* ---------------------------------------------------------------------- */
int or_sim_f64_begin(or_sim *s) {
    free(s->electric64);
    free(s->magnetic64);
    s->electric64 = malloc(sizeof(double) * 3 * s->cells);
    s->magnetic64 = malloc(sizeof(double) * 3 * s->cells);
    if (!s->electric64 || !s->magnetic64) return fail("out of host memory for the f64 track");
    for (uint64_t i = 0; i < 3 * s->cells; ++i) {
        s->electric64[i] = (double)s->electric[i];
        s->magnetic64[i] = (double)s->magnetic[i];
    }
    s->step64 = s->step;
    return 0;
}

double *or_sim_electric_f64(or_sim *s) { return s->electric64; }
double *or_sim_magnetic_f64(or_sim *s) { return s->magnetic64; }
uint64_t or_sim_step_count_f64(const or_sim *s) { return s->step64; }

static void sweep_magnetic64(or_sim *s) {
    const uint32_t *extent = s->grid.extent;
    const uint64_t strides[3] = {1, extent[0], (uint64_t)extent[0] * extent[1]};
    int threads = thread_count(s);
    for (int a = 0; a < 3; ++a) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        const float *gain_b = s->magnetic_gains[b], *gain_c = s->magnetic_gains[c];
        double *target = s->magnetic64 + (uint64_t)a * s->cells;
        const double *field_b = s->electric64 + (uint64_t)b * s->cells;
        const double *field_c = s->electric64 + (uint64_t)c * s->cells;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t zz = 0; zz < (int64_t)extent[2]; ++zz) {
            uint32_t z = (uint32_t)zz;
            for (uint32_t y = 0; y < extent[1]; ++y)
                for (uint32_t x = 0; x < extent[0]; ++x) {
                    uint32_t coord[3] = {x, y, z};
                    uint64_t i = ((uint64_t)z * extent[1] + y) * extent[0] + x;
                    /* past a high wall the neighbour reads as zero (src/cpu.rs:335-374) */
                    double next_c = coord[b] + 1 < extent[b] ? field_c[i + strides[b]] : 0.0;
                    double next_b = coord[c] + 1 < extent[c] ? field_b[i + strides[c]] : 0.0;
                    double curl = (double)gain_b[coord[b]] * (next_c - field_c[i]) -
                                  (double)gain_c[coord[c]] * (next_b - field_b[i]);
                    float absorb_row;
                    const float *absorb_x = magnetic_row(&s->absorber, a, y, z, &absorb_row);
                    coeff_t k = s->coefficients[s->material_index[i]];
                    double loss = (double)k.magnetic_loss + (double)absorb_row + (double)absorb_x[x];
                    target[i] = ((1.0 - loss) * target[i] - (double)k.magnetic_gain * curl) / (1.0 + loss);
                }
        }
    }
}

static void sweep_electric64(or_sim *s) {
    const uint32_t *extent = s->grid.extent;
    const uint64_t strides[3] = {1, extent[0], (uint64_t)extent[0] * extent[1]};
    int threads = thread_count(s);
    for (int a = 0; a < 3; ++a) {
        int b = (a + 1) % 3, c = (a + 2) % 3;
        const float *gain_b = s->electric_gains[b], *gain_c = s->electric_gains[c];
        double *target = s->electric64 + (uint64_t)a * s->cells;
        const double *field_b = s->magnetic64 + (uint64_t)b * s->cells;
        const double *field_c = s->magnetic64 + (uint64_t)c * s->cells;
#pragma omp parallel for schedule(static) num_threads(threads)
        for (int64_t zz = 0; zz < (int64_t)extent[2]; ++zz) {
            uint32_t z = (uint32_t)zz;
            for (uint32_t y = 0; y < extent[1]; ++y)
                for (uint32_t x = 0; x < extent[0]; ++x) {
                    uint32_t coord[3] = {x, y, z};
                    if (coord[b] == 0 || coord[c] == 0) continue; /* low-wall PEC */
                    uint64_t i = ((uint64_t)z * extent[1] + y) * extent[0] + x;
                    double curl = (double)gain_b[coord[b]] * (field_c[i] - field_c[i - strides[b]]) -
                                  (double)gain_c[coord[c]] * (field_b[i] - field_b[i - strides[c]]);
                    float absorb_row;
                    const float *absorb_x = electric_row(&s->absorber, a, y, z, &absorb_row);
                    coeff_t k = s->coefficients[s->material_index[i]];
                    double loss = (double)k.electric_loss + (double)absorb_row + (double)absorb_x[x];
                    target[i] = ((1.0 - loss) * target[i] + (double)k.electric_gain * curl) / (1.0 + loss);
                }
        }
    }
}

static void apply_sources64(or_sim *s, float time, int magnetic) {
    const uint32_t *extent = s->grid.extent;
    for (uint32_t k = 0; k < s->source_count; ++k) {
        dp_injection inj[2];
        source_injections(&s->sources[k], &s->grid, time, inj);
        for (int j = 0; j < 2; ++j) {
            const dp_injection *q = &inj[j];
            if (q->value == 0.0f || (int)q->magnetic != magnetic) continue;
            double *target = (q->magnetic ? s->magnetic64 : s->electric64) + (uint64_t)q->component * s->cells;
            for (uint32_t dz = 0; dz < q->extent[2]; ++dz)
                for (uint32_t dy = 0; dy < q->extent[1]; ++dy)
                    for (uint32_t dx = 0; dx < q->extent[0]; ++dx) {
                        uint32_t c[3] = {q->origin[0] + dx, q->origin[1] + dy, q->origin[2] + dz};
                        float p[3];
                        grid_cell_center(&s->grid, c, p);
                        uint64_t i = ((uint64_t)c[2] * extent[1] + c[1]) * extent[0] + c[0];
                        target[i] += (double)q->value * (double)injection_weight(q, p);
                    }
        }
    }
}

int or_sim_f64_advance(or_sim *s, uint64_t steps) {
    if (!s->electric64) return fail("the f64 track was not started");
    for (uint64_t n = 0; n < steps; ++n) {
        float time = (float)(s->step64 + 1) * grid_time_step(&s->grid);
        sweep_magnetic64(s);
        apply_sources64(s, time, 1);
        sweep_electric64(s);
        apply_sources64(s, time, 0);
        s->step64 += 1;
    }
    return 0;
}
