/*
 * FP64 oracle of one substep (test infrastructure only; see sph_oracle.h).  The equations
 * are those of sph_ref32.c evaluated in binary64; every integer decision (sorted order,
 * start/end table, the cells each particle visits) comes from the FP32 run, so the
 * difference between the two measures the precision of the FP32 arithmetic only.
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>

#define PI64 3.14159265359 /* the shaders' PI constant, in double */

static void clamp64(const sphref_params *p, double margin, double *pos, double *vel) {
    for (int a = 0; a < 2; ++a) {
        if (pos[a] < (double)p->min_xyz[a] + margin) {
            pos[a] = (double)p->min_xyz[a] + margin;
            vel[a] = -vel[a];
        } else if (pos[a] > (double)p->max_xyz[a] - margin) {
            pos[a] = (double)p->max_xyz[a] - margin;
            vel[a] = -vel[a];
        }
    }
    if (pos[2] < (double)p->min_xyz[2] + margin) {
        pos[2] = (double)p->min_xyz[2] + margin;
        vel[2] = -vel[2] * 0.95;
    } else if (pos[2] > (double)p->max_xyz[2] - margin) {
        pos[2] = (double)p->max_xyz[2] - margin;
        vel[2] = -vel[2];
    }
}

static void kick_drift(const sphref_params *p, double margin, double *pos, double *vel) {
    const double dt = p->dt;
    vel[2] += -(double)p->g * dt;
    for (int a = 0; a < 3; ++a) pos[a] += vel[a] * dt;
    clamp64(p, margin, pos, vel);
}

static inline uint32_t key_of_cell(uint32_t mask, int32_t x, int32_t y, int32_t z) {
    return sphref_cell_hash((uint32_t)x, (uint32_t)y, (uint32_t)z) & mask;
}

int sphref64_substep(const sphref_params *p, uint32_t n, const float *pos4, const float *vel4,
                     const uint32_t *sorted_idx, const int32_t *start, const int32_t *end,
                     const float *pred_pos4_f32, double *density, double *near_density,
                     double *vel_force3, double *pos_out3, double *vel_out3, double *dv_scale) {
    sphref_derived d;
    sphref_derive(p, &d);
    const double R = p->smoothing_radius, m = p->particle_mass, rho0 = p->target_density;
    const double k = p->pressure_multiplier, kn = p->near_pressure_multiplier;
    const double Rf = p->smoothing_radius;
    const double slope_scale = 15.0 / (PI64 * pow(R, 5));
    const double near_slope_scale = -45.0 / (PI64 * pow(R, 6));
    const double near_vol_inv = 15.0 / (PI64 * pow(R, 6));
    double *ps = malloc((size_t)n * 24), *vs = malloc((size_t)n * 24);
    double *pps = malloc((size_t)n * 24), *pvs = malloc((size_t)n * 24);
    if (!ps || !vs || !pps || !pvs) {
        free(ps); free(vs); free(pps); free(pvs);
        return -1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        size_t i = sorted_idx[j];
        for (int a = 0; a < 3; ++a) {
            ps[3 * j + a] = pos4[4 * i + a];
            vs[3 * j + a] = vel4[4 * i + a];
            pps[3 * j + a] = ps[3 * j + a];
            pvs[3 * j + a] = vs[3 * j + a];
        }
        kick_drift(p, (double)p->particle_radius, &pps[3 * j], &pvs[3 * j]);
    }

    /* density: cells chosen from the FP32 predicted position (calculateDensities.comp:117) */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float *s32 = &pred_pos4_f32[4 * j];
        const double *s = &pps[3 * j];
        double M = 0.0, Mn = 0.0;
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    float q[3] = {s32[0] + p->smoothing_radius * (float)dx,
                                  s32[1] + p->smoothing_radius * (float)dy,
                                  s32[2] + p->smoothing_radius * (float)dz};
                    int out = 0;
                    for (int a = 0; a < 3; ++a)
                        if (q[a] < p->min_xyz[a] || q[a] > p->max_xyz[a]) out = 1;
                    if (out) continue;
                    int32_t c[3];
                    for (int a = 0; a < 3; ++a)
                        c[a] = (int32_t)floorf((q[a] - p->min_xyz[a]) / p->smoothing_radius);
                    uint32_t key = key_of_cell(d.mask, c[0], c[1], c[2]);
                    if (start[key] == -1) continue;
                    for (int32_t i = start[key]; i < end[key]; ++i) {
                        const double *pi = &pps[3 * (int64_t)i];
                        double dist = sqrt((pi[0] - s[0]) * (pi[0] - s[0]) +
                                           (pi[1] - s[1]) * (pi[1] - s[1]) +
                                           (pi[2] - s[2]) * (pi[2] - s[2]));
                        if (dist >= Rf) continue;
                        M += m * (R - dist) * (R - dist);
                        Mn += m * (R - dist) * (R - dist) * (R - dist);
                    }
                }
        density[j] = fmax(0.01, M / (double)p->volume);
        near_density[j] = fmax(0.01, Mn * near_vol_inv);
    }

    /* force (Jacobi) and integrate */
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t j = 0; j < (int64_t)n; ++j) {
        const float *s32 = &pred_pos4_f32[4 * j];
        const double *s = &pps[3 * j], *u = &pvs[3 * j];
        double Fp[3] = {0, 0, 0}, Fv[3] = {0, 0, 0}, mag = 0.0;
        const double myd = density[j], near_p = near_density[j] * kn;
        int32_t c0[3];
        int outside = 0;
        for (int a = 0; a < 3; ++a)
            if (s32[a] < p->min_xyz[a] || s32[a] > p->max_xyz[a]) outside = 1;
        for (int a = 0; a < 3; ++a)
            c0[a] = outside ? -1
                            : (int32_t)floorf((s32[a] - p->min_xyz[a]) / p->smoothing_radius);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    uint32_t key = key_of_cell(d.mask, c0[0] + dx, c0[1] + dy, c0[2] + dz);
                    if (start[key] == -1) continue;
                    for (int32_t i = start[key]; i < end[key]; ++i) {
                        if ((int64_t)i == j) continue;
                        const double *pi = &pps[3 * (int64_t)i];
                        double dv[3] = {pi[0] - s[0], pi[1] - s[1], pi[2] - s[2]};
                        double dist2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
                        if (dist2 > R * R) continue;
                        double dist = sqrt(dist2);
                        if (dist < 1e-7) continue;
                        double slope = dist >= R ? 0.0 : (dist - R) * slope_scale;
                        double rho_i = density[i];
                        double shared = ((rho_i - rho0) * k + (myd - rho0) * k) / 2.0;
                        double a = shared * slope * m / rho_i;
                        double infl = dist >= R ? 0.0 : (R - dist) * (R - dist);
                        double b = m * infl / rho_i;
                        double nrho_i = near_density[i];
                        double nslope = dist >= R ? 0.0 : (R - dist) * (R - dist) * near_slope_scale;
                        double c = (near_p + nrho_i * kn) / 2.0 * nslope * m / nrho_i;
                        double du2 = 0.0;
                        for (int q = 0; q < 3; ++q) {
                            Fp[q] += (a + c) * (dv[q] / dist);
                            Fv[q] += b * (vs[3 * (int64_t)i + q] - u[q]);
                            du2 += (vs[3 * (int64_t)i + q] - u[q]) * (vs[3 * (int64_t)i + q] - u[q]);
                        }
                        mag += fabs(a + c) + fabs(b) * sqrt(du2);
                    }
                }
        double v[3], x[3];
        for (int q = 0; q < 3; ++q) {
            v[q] = vs[3 * j + q] + (Fp[q] + Fv[q]) / myd * (double)p->dt;
            vel_force3[3 * j + q] = v[q];
            x[q] = ps[3 * j + q];
        }
        if (dv_scale) dv_scale[j] = mag / myd * (double)p->dt;
        kick_drift(p, (double)d.integrate_margin, x, v);
        for (int q = 0; q < 3; ++q) {
            pos_out3[3 * j + q] = x[q];
            vel_out3[3 * j + q] = v[q];
        }
    }
    free(ps); free(vs); free(pps); free(pvs);
    return 0;
}
