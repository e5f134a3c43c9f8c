/*
 * FP32 restatement of the reference's per-substep SPH chain (test infrastructure only; see
 * sph_oracle.h).  Build flags: -ffp-contract=off, no -ffast-math, so every expression below
 * is evaluated as written, each operation rounded to binary32.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SPHREF_PI 3.14159265359f /* calculateDensities.comp:75 */

void sphref_derive(const sphref_params *p, sphref_derived *d) {
    const float R = p->smoothing_radius;
    d->mask = p->table_cells - 1u;
    d->R2 = R * R;
    d->slope_scale = 15.0f / (SPHREF_PI * powf(R, 5.0f));
    d->near_slope_scale = -45.0f / (SPHREF_PI * powf(R, 6.0f));
    d->near_vol_inv = 15.0f / (SPHREF_PI * powf(R, 6.0f));
    d->integrate_margin = p->particle_radius + 0.5f;
}

/* cellCoordToHash: updateCellEntries.comp:36-52 (identical copies in the other shaders). */
uint32_t sphref_cell_hash(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t h = (x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

/* floor((v - lo) / R) converted to a 32-bit cell coordinate.  Inside the box the value is
 * non-negative; outside it wraps through int32 like the force shader's ivec3->uvec3 cast. */
static inline uint32_t cell_coord(float v, float lo, float R) {
    return (uint32_t)(int32_t)floorf((v - lo) / R);
}

/* Wall clamp shared by predictPositions.comp:81-113 and updateVelocities.comp:83-115 (the
 * latter with margin particleRadius + 0.5).  vel *= 1.00 is exact and omitted. */
static inline void wall_clamp(const sphref_params *p, float margin, float *pos, float *vel) {
    for (int a = 0; a < 2; ++a) {
        if (pos[a] < p->min_xyz[a] + margin) {
            pos[a] = p->min_xyz[a] + margin;
            vel[a] = vel[a] * -1.0f;
        } else if (pos[a] > p->max_xyz[a] - margin) {
            pos[a] = p->max_xyz[a] - margin;
            vel[a] = vel[a] * -1.0f;
        }
    }
    if (pos[2] < p->min_xyz[2] + margin) {
        pos[2] = p->min_xyz[2] + margin;
        vel[2] = vel[2] * -1.0f;
        vel[2] = vel[2] * 0.95f;
    } else if (pos[2] > p->max_xyz[2] - margin) {
        pos[2] = p->max_xyz[2] - margin;
        vel[2] = vel[2] * -1.0f;
    }
}

/* predictPositions.comp:69-117 */
void sphref_predict(const sphref_params *p, uint32_t n, const float *pos4, const float *vel4,
                    float *pred_pos4, float *pred_vel4) {
    const float dt = p->dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float pos[3] = {pos4[4 * i + 0], pos4[4 * i + 1], pos4[4 * i + 2]};
        float vel[3] = {vel4[4 * i + 0], vel4[4 * i + 1], vel4[4 * i + 2]};
        /* vel += vec3(0, 0, -g) * dt */
        vel[0] = vel[0] + 0.0f * dt;
        vel[1] = vel[1] + 0.0f * dt;
        vel[2] = vel[2] + (-p->g) * dt;
        /* pos += vel * dt */
        pos[0] = pos[0] + vel[0] * dt;
        pos[1] = pos[1] + vel[1] * dt;
        pos[2] = pos[2] + vel[2] * dt;
        wall_clamp(p, p->particle_radius, pos, vel);
        for (int a = 0; a < 3; ++a) {
            pred_pos4[4 * i + a] = pos[a];
            pred_vel4[4 * i + a] = vel[a];
        }
        pred_pos4[4 * i + 3] = 1.0f;
        pred_vel4[4 * i + 3] = 1.0f;
    }
}

/* updateCellEntries.comp:69-80: key of the CURRENT position; value = particle index. */
void sphref_hash(const sphref_params *p, uint32_t n, const float *pos4, uint32_t *keys) {
    sphref_derived d;
    sphref_derive(p, &d);
    const float R = p->smoothing_radius;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        uint32_t cx = cell_coord(pos4[4 * i + 0], p->min_xyz[0], R);
        uint32_t cy = cell_coord(pos4[4 * i + 1], p->min_xyz[1], R);
        uint32_t cz = cell_coord(pos4[4 * i + 2], p->min_xyz[2], R);
        keys[i] = sphref_cell_hash(cx, cy, cz) & d.mask;
    }
}

/* The reference sorts CellEntry{key, index} with a 4 x 8-bit LSD radix sort whose scatter is
 * stable (multiRadixSort.comp:54-146 driven by commands.c:106-161), so its output is the
 * pairs ordered by (key, index).  Restated here as a stable counting sort over the table.
 * (For keys >= 2^24 the reference builds its table from the pass-2 buffer instead, which
 * has no meaningful semantics; the oracle always uses the full sort.) */
int sphref_sort(uint32_t n, uint32_t table_cells, const uint32_t *keys, uint32_t *out_keys,
                uint32_t *out_idx) {
    uint32_t *count = (uint32_t *)calloc((size_t)table_cells + 1u, sizeof(uint32_t));
    if (!count) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        if (keys[i] >= table_cells) { free(count); return -1; }
        count[keys[i] + 1u]++;
    }
    for (uint32_t k = 0; k < table_cells; ++k) count[k + 1u] += count[k];
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t slot = count[keys[i]]++;
        out_keys[slot] = keys[i];
        out_idx[slot] = i;
    }
    free(count);
    return 0;
}

/* clearStartIndices.comp:47-52 then updateStartIndices.comp:40-64. */
void sphref_table(uint32_t n, uint32_t table_cells, const uint32_t *sorted_keys,
                  int32_t *start, int32_t *end) {
    for (uint32_t k = 0; k < table_cells; ++k) {
        start[k] = -1;
        end[k] = -1;
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t cur = sorted_keys[i];
        if (i == 0) {
            start[cur] = 0;
        } else {
            uint32_t prev = sorted_keys[i - 1];
            if (cur != prev) {
                start[cur] = (int32_t)i;
                end[prev] = (int32_t)i;
            }
        }
        if (i == n - 1u) end[cur] = (int32_t)n;
    }
}

/* samplePointToKey (calculateDensities.comp:69-73): ~0u when the point is outside the box. */
static inline uint32_t sample_point_key(const sphref_params *p, uint32_t mask, const float q[3]) {
    for (int a = 0; a < 3; ++a)
        if (q[a] < p->min_xyz[a]) return 0xFFFFFFFFu;
    for (int a = 0; a < 3; ++a)
        if (q[a] > p->max_xyz[a]) return 0xFFFFFFFFu;
    const float R = p->smoothing_radius;
    return sphref_cell_hash(cell_coord(q[0], p->min_xyz[0], R), cell_coord(q[1], p->min_xyz[1], R),
                            cell_coord(q[2], p->min_xyz[2], R)) &
           mask;
}

/* The 27 keys the density pass visits for sample point s, in the shader's (dz, dy, dx)
 * order (calculateDensities.comp:114-118).  ~0u marks a skipped (out-of-box) cell. */
static void density_keys(const sphref_params *p, uint32_t mask, const float s[3], uint32_t out[27]) {
    const float R = p->smoothing_radius;
    int c = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                /* samplePoint + smoothingRadius * vec3(dx, dy, dz) */
                float q[3] = {s[0] + R * (float)dx, s[1] + R * (float)dy, s[2] + R * (float)dz};
                out[c++] = sample_point_key(p, mask, q);
            }
}

/* distance(a, b) = length(a - b) = sqrt(dot(a - b, a - b)), dot evaluated left to right. */
static inline float dist3(const float *a, const float *b) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

static inline void density_accumulate(const sphref_params *p, float dist, float *M, float *Mn) {
    const float R = p->smoothing_radius;
    const float m = p->particle_mass;
    /* smoothingKernel / nearDensitySmoothingKernel return 0 for dist >= R, and
     * M + m*0 == M for M >= +0, so the out-of-range term is skipped. */
    if (dist >= R) return;
    float infl = (R - dist) * (R - dist);
    float ninfl = (R - dist) * (R - dist) * (R - dist);
    *M = *M + m * infl;
    *Mn = *Mn + m * ninfl;
}

static inline void density_finish(const sphref_params *p, const sphref_derived *d, float M,
                                  float Mn, float *rho, float *rho_n) {
    float density = M / p->volume;                 /* calculateDensities.comp:133 */
    float near_density = Mn * d->near_vol_inv;     /* :134 */
    *rho = fmaxf(0.01f, density);                  /* :151 */
    *rho_n = fmaxf(0.01f, near_density);           /* :152 */
}

/* calculateDensities.comp:96-155 */
void sphref_density(const sphref_params *p, uint32_t n, const float *pred_pos4,
                    const int32_t *start, const int32_t *end, float *density,
                    float *near_density) {
    sphref_derived d;
    sphref_derive(p, &d);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float *s = &pred_pos4[4 * j];
        uint32_t keys[27];
        density_keys(p, d.mask, s, keys);
        float M = 0.0f, Mn = 0.0f;
        for (int c = 0; c < 27; ++c) {
            if (keys[c] == 0xFFFFFFFFu) continue;
            int32_t st = start[keys[c]];
            if (st == -1) continue;
            int32_t en = end[keys[c]];
            for (int32_t i = st; i < en; ++i)
                density_accumulate(p, dist3(&pred_pos4[4 * (int64_t)i], s), &M, &Mn);
        }
        density_finish(p, &d, M, Mn, &density[j], &near_density[j]);
    }
}

void sphref_density_bruteforce(const sphref_params *p, uint32_t n, const float *pred_pos4,
                               const uint32_t *sorted_keys, float *density,
                               float *near_density) {
    sphref_derived d;
    sphref_derive(p, &d);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float *s = &pred_pos4[4 * j];
        uint32_t keys[27];
        density_keys(p, d.mask, s, keys);
        float M = 0.0f, Mn = 0.0f;
        for (int c = 0; c < 27; ++c) {
            if (keys[c] == 0xFFFFFFFFu) continue;
            /* every particle whose key matches, in ascending sorted position */
            for (uint32_t i = 0; i < n; ++i)
                if (sorted_keys[i] == keys[c])
                    density_accumulate(p, dist3(&pred_pos4[4 * (int64_t)i], s), &M, &Mn);
        }
        density_finish(p, &d, M, Mn, &density[j], &near_density[j]);
    }
}

/* applyPressureAndViscosityForce.comp:275-337, with neighbours' velocities read from the
 * pre-force buffer vel4 and the result written to vel_out4. */
void sphref_force(const sphref_params *p, uint32_t n, const float *pred_pos4,
                  const float *pred_vel4, const float *vel4, const float *density,
                  const float *near_density, const int32_t *start, const int32_t *end,
                  float *vel_out4) {
    sphref_derived d;
    sphref_derive(p, &d);
    const float R = p->smoothing_radius;
    const float m = p->particle_mass;
    const float rho0 = p->target_density;
    const float k = p->pressure_multiplier;
    const float kn = p->near_pressure_multiplier;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float *s = &pred_pos4[4 * j];
        const float *u = &pred_vel4[4 * j];
        float Fp[3] = {0.0f, 0.0f, 0.0f};
        float Fv[3] = {0.0f, 0.0f, 0.0f};
        const float my_density = density[j];
        const float near_pressure = near_density[j] * kn;
        const float my_pressure_term = (my_density - rho0) * k;
        int32_t c0[3];
        int outside = 0;
        for (int a = 0; a < 3; ++a)
            if (s[a] < p->min_xyz[a]) outside = 1;
        for (int a = 0; a < 3; ++a)
            if (s[a] > p->max_xyz[a]) outside = 1;
        for (int a = 0; a < 3; ++a)
            c0[a] = outside ? -1 : (int32_t)floorf((s[a] - p->min_xyz[a]) / R);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    uint32_t key = sphref_cell_hash((uint32_t)(c0[0] + dx), (uint32_t)(c0[1] + dy),
                                                    (uint32_t)(c0[2] + dz)) &
                                   d.mask;
                    int32_t st = start[key];
                    if (st == -1) continue;
                    int32_t en = end[key];
                    for (int32_t i = st; i < en; ++i) {
                        if ((int64_t)i == j) continue;
                        const float *pi = &pred_pos4[4 * (int64_t)i];
                        float dv[3] = {pi[0] - s[0], pi[1] - s[1], pi[2] - s[2]};
                        float dist2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
                        if (dist2 > d.R2) continue;
                        float dist = sqrtf(dist2);
                        if (fabsf(dist) < 1e-7f) continue;
                        float dir[3] = {dv[0] / dist, dv[1] / dist, dv[2] / dist};
                        float slope = dist >= R ? 0.0f : (dist - R) * d.slope_scale;
                        float rho_i = density[i];
                        /* calculateSharedPressure(density, myDensity) */
                        float shared = ((rho_i - rho0) * k + my_pressure_term) / 2.0f;
                        float a = shared * slope * m / rho_i;
                        for (int q = 0; q < 3; ++q) Fp[q] = Fp[q] + a * dir[q];
                        float infl = dist >= R ? 0.0f : (R - dist) * (R - dist);
                        float b = m * infl / density[i];
                        const float *vi = &vel4[4 * (int64_t)i];
                        for (int q = 0; q < 3; ++q) Fv[q] = Fv[q] + b * (vi[q] - u[q]);
                        float nrho_i = near_density[i];
                        float shared_near = (near_pressure + nrho_i * kn) / 2.0f;
                        float x = R - dist;
                        float nslope = dist >= R ? 0.0f : x * x * d.near_slope_scale;
                        float c = shared_near * nslope * m / nrho_i;
                        for (int q = 0; q < 3; ++q) Fp[q] = Fp[q] + c * dir[q];
                    }
                }
        /* velocities[j] += vec4((pressureForce + viscosityForce + nearPressureForce)
         *                       / densities[j] * dt, 1), nearPressureForce == 0 */
        for (int q = 0; q < 3; ++q)
            vel_out4[4 * j + q] = vel4[4 * j + q] + (Fp[q] + Fv[q] + 0.0f) / my_density * p->dt;
        vel_out4[4 * j + 3] = vel4[4 * j + 3] + 1.0f;
    }
}

/* updateVelocities.comp:71-119 (colour handled by sphref_color). */
void sphref_integrate(const sphref_params *p, uint32_t n, float *pos4, float *vel4) {
    sphref_derived d;
    sphref_derive(p, &d);
    const float dt = p->dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float pos[3] = {pos4[4 * i + 0], pos4[4 * i + 1], pos4[4 * i + 2]};
        float vel[3] = {vel4[4 * i + 0], vel4[4 * i + 1], vel4[4 * i + 2]};
        vel[0] = vel[0] + 0.0f * dt;
        vel[1] = vel[1] + 0.0f * dt;
        vel[2] = vel[2] + (-p->g) * dt;
        /* pos += vel * dt * dt_scale, dt_scale = 1.0 */
        for (int a = 0; a < 3; ++a) pos[a] = pos[a] + vel[a] * dt * 1.0f;
        wall_clamp(p, d.integrate_margin, pos, vel);
        for (int a = 0; a < 3; ++a) {
            pos4[4 * i + a] = pos[a];
            vel4[4 * i + a] = vel[a];
        }
        pos4[4 * i + 3] = 1.0f;
        vel4[4 * i + 3] = 1.0f;
    }
}

/* updateVelocities.comp:120-141: blue -> cyan -> yellow -> red ramp of |v| / 60. */
void sphref_color(uint32_t n, const float *vel4, float *color4) {
    static const float blue[3] = {0.3f, 0.3f, 1.0f}, cyan[3] = {0.0f, 1.0f, 1.0f},
                       yellow[3] = {1.0f, 1.0f, 0.0f}, red[3] = {1.0f, 0.3f, 0.3f};
    for (uint32_t i = 0; i < n; ++i) {
        const float *v = &vel4[4 * (size_t)i];
        float speed = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        float t = fminf(fmaxf(speed / 60.0f, 0.0f), 1.0f);
        const float *a, *b;
        float w;
        if (t < 0.33f) { a = blue; b = cyan; w = t / 0.33f; }
        else if (t < 0.66f) { a = cyan; b = yellow; w = (t - 0.33f) / 0.33f; }
        else { a = yellow; b = red; w = (t - 0.66f) / 0.34f; }
        for (int q = 0; q < 3; ++q) color4[4 * (size_t)i + q] = a[q] * (1.0f - w) + b[q] * w;
        color4[4 * (size_t)i + 3] = 1.0f;
    }
}

/* Voxel counts of the density map: initializeDensityMap (renderer.c:774-783) scales the
 * box extent by the float reciprocal of the voxel edge and takes ceil. */
void sphref_density_map_dims(const sphref_params *p, float voxel_size, uint32_t dims[3]) {
    const float inv = 1.0f / voxel_size;
    for (int a = 0; a < 3; ++a) dims[a] = (uint32_t)(int)ceil((double)((p->max_xyz[a] - p->min_xyz[a]) * inv));
}

/* outOfBounds (updateDensityMap.comp:94-97): against the voxel count the shader derives,
 * ceil((max - min) / voxel_size); used for voxels and, as the shader does, for cells. */
static int map_out_of_bounds(const int32_t v[3], const int32_t mx[3]) {
    if (v[0] < 0 || v[1] < 0 || v[2] < 0) return 1;
    if (v[0] > mx[0] || v[1] > mx[1] || v[2] > mx[2]) return 1;
    return v[0] == mx[0] && v[1] == mx[1] && v[2] == mx[2];
}

/* updateDensityMap.comp:114-148 over the state and table a substep left (positions in
 * sorted order, startIndices/endIndices of that substep).  Voxels the shader does not
 * store (only the corner voxel, and only when the two voxel-count formulas disagree) are
 * written as 0. */
void sphref_density_map(const sphref_params *p, uint32_t n, const float *pos4,
                        const int32_t *start, const int32_t *end, float voxel_size,
                        float *out) {
    (void)n;
    const float R = p->smoothing_radius;
    const uint32_t mask = p->table_cells - 1u;
    uint32_t dims[3];
    sphref_density_map_dims(p, voxel_size, dims);
    int32_t mx[3];
    for (int a = 0; a < 3; ++a) mx[a] = (int32_t)ceilf((p->max_xyz[a] - p->min_xyz[a]) / voxel_size);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t vz = 0; vz < (int64_t)dims[2]; ++vz)
        for (uint32_t vy = 0; vy < dims[1]; ++vy)
            for (uint32_t vx = 0; vx < dims[0]; ++vx) {
                const size_t o = (size_t)vx + (size_t)dims[0] * ((size_t)vy + (size_t)dims[1] * (size_t)vz);
                const int32_t v[3] = {(int32_t)vx, (int32_t)vy, (int32_t)vz};
                if (map_out_of_bounds(v, mx)) {
                    out[o] = 0.0f;
                    continue;
                }
                float c[3];
                for (int a = 0; a < 3; ++a) c[a] = p->min_xyz[a] + ((float)v[a] + 0.5f) * voxel_size;
                int32_t vc[3];
                const int outside = c[0] < p->min_xyz[0] || c[1] < p->min_xyz[1] || c[2] < p->min_xyz[2] ||
                                    c[0] > p->max_xyz[0] || c[1] > p->max_xyz[1] || c[2] > p->max_xyz[2];
                for (int a = 0; a < 3; ++a)
                    vc[a] = outside ? -1 : (int32_t)floorf((c[a] - p->min_xyz[a]) / R);
                float density = 0.0f;
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int32_t cell[3] = {vc[0] + dx, vc[1] + dy, vc[2] + dz};
                            if (map_out_of_bounds(cell, mx)) continue;
                            const uint32_t key = sphref_cell_hash((uint32_t)cell[0], (uint32_t)cell[1],
                                                                  (uint32_t)cell[2]) & mask;
                            const int32_t s0 = start[key];
                            if (s0 == -1) continue;
                            const int32_t e0 = end[key];
                            for (int32_t i = s0; i < e0; ++i) {
                                const float *q = &pos4[4 * (size_t)i];
                                const float ex = c[0] - q[0], ey = c[1] - q[1], ez = c[2] - q[2];
                                const float d2 = ex * ex + ey * ey + ez * ez;
                                if (d2 < R * R) {
                                    const float dist = sqrtf(d2);
                                    const float k = dist >= R ? 0.0f : (R - dist) * (R - dist);
                                    density = density + p->particle_mass * k;
                                }
                            }
                        }
                density = density * 1.0f;
                out[o] = density / p->volume;
            }
}

static void gather4(uint32_t n, const uint32_t *idx, const float *src, float *dst) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j) memcpy(&dst[4 * j], &src[4 * (size_t)idx[j]], 16);
}

/* One substep in the order of commands.c:193-402.  The reference keeps the predicted
 * arrays beside the state and permutes all of them in scatterPass.comp; gathering the
 * predicted arrays is equivalent to predicting from the gathered state because predict is
 * a pure per-particle function. */
int sphref_substep(const sphref_params *p, uint32_t n, float *pos4, float *vel4,
                   const sphref_substep_out *out) {
    const uint32_t T = p->table_cells;
    if (n == 0 || T == 0) return -1;
    int rc = -1;
    float *pp = malloc((size_t)n * 16), *pv = malloc((size_t)n * 16);
    float *ps = malloc((size_t)n * 16), *vs = malloc((size_t)n * 16);
    float *pps = malloc((size_t)n * 16), *pvs = malloc((size_t)n * 16);
    float *vf = malloc((size_t)n * 16);
    float *rho = malloc((size_t)n * 4), *rhon = malloc((size_t)n * 4);
    uint32_t *keys = malloc((size_t)n * 4), *skeys = malloc((size_t)n * 4);
    uint32_t *sidx = malloc((size_t)n * 4);
    int32_t *st = malloc((size_t)T * 4), *en = malloc((size_t)T * 4);
    if (!pp || !pv || !ps || !vs || !pps || !pvs || !vf || !rho || !rhon || !keys || !skeys ||
        !sidx || !st || !en)
        goto done;

    sphref_predict(p, n, pos4, vel4, pp, pv);
    sphref_hash(p, n, pos4, keys);
    if (sphref_sort(n, T, keys, skeys, sidx) != 0) goto done;
    gather4(n, sidx, pos4, ps);
    gather4(n, sidx, vel4, vs);
    gather4(n, sidx, pp, pps);
    gather4(n, sidx, pv, pvs);
    sphref_table(n, T, skeys, st, en);
    sphref_density(p, n, pps, st, en, rho, rhon);
    sphref_force(p, n, pps, pvs, vs, rho, rhon, st, en, vf);

    if (out) {
        if (out->cell_keys) memcpy(out->cell_keys, keys, (size_t)n * 4);
        if (out->sorted_keys) memcpy(out->sorted_keys, skeys, (size_t)n * 4);
        if (out->sorted_idx) memcpy(out->sorted_idx, sidx, (size_t)n * 4);
        if (out->start) memcpy(out->start, st, (size_t)T * 4);
        if (out->end) memcpy(out->end, en, (size_t)T * 4);
        if (out->pred_pos) memcpy(out->pred_pos, pps, (size_t)n * 16);
        if (out->pred_vel) memcpy(out->pred_vel, pvs, (size_t)n * 16);
        if (out->density) memcpy(out->density, rho, (size_t)n * 4);
        if (out->near_density) memcpy(out->near_density, rhon, (size_t)n * 4);
        if (out->vel_force) memcpy(out->vel_force, vf, (size_t)n * 16);
    }

    sphref_integrate(p, n, ps, vf);
    memcpy(pos4, ps, (size_t)n * 16);
    memcpy(vel4, vf, (size_t)n * 16);
    rc = 0;
done:
    free(pp); free(pv); free(ps); free(vs); free(pps); free(pvs); free(vf);
    free(rho); free(rhon); free(keys); free(skeys); free(sidx); free(st); free(en);
    return rc;
}

int sphref_step(const sphref_params *p, uint32_t n, float *pos4, float *vel4, uint32_t frames) {
    for (uint32_t f = 0; f < frames; ++f)
        for (uint32_t s = 0; s < p->substeps; ++s)
            if (sphref_substep(p, n, pos4, vel4, NULL) != 0) return -1;
    return 0;
}
