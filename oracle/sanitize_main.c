/*
 * Sanitizer driver for the CPU oracle (test infrastructure only).
 *
 * Runs every entry point of sph_ref32.c and sph_ref64.c on small scenes chosen to reach
 * their edge paths: hash collisions (T much smaller than the cell count), a single table
 * cell (T = 1, every particle in one run), one particle, particles on the walls, a box
 * with a negative and non-integer minimum, the near-pressure term on, several substeps,
 * the density map with voxel centres outside the box.  Built by `make -C oracle sanitize`
 * twice: with AddressSanitizer + UndefinedBehaviorSanitizer (serial, every report fatal),
 * and with ThreadSanitizer over the OpenMP loops (LLVM's OpenMP runtime with its Archer
 * tool, which tells ThreadSanitizer about the runtime's barriers; libgomp does not, and
 * every implicit barrier then reads as a race).  tests/test_oracle_sanitized.py runs both.  Exit status 0 and "sanitize: ok" mean every call returned and no sanitizer
 * reported.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

static uint32_t lcg_state = 12345u;
static float frand(void) {
    lcg_state = lcg_state * 1664525u + 1013904223u;
    return (float)(lcg_state >> 8) * (1.0f / 16777216.0f);
}

static void fill(uint32_t n, const float lo[3], const float hi[3], float *pos, float *vel) {
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            pos[4 * i + a] = lo[a] + (hi[a] - lo[a]) * frand();
            vel[4 * i + a] = (frand() - 0.5f) * 0.2f;
        }
        pos[4 * i + 3] = 0.0f;
        vel[4 * i + 3] = 0.0f;
    }
}

static sphref_params scene(const float mn[3], const float mx[3], uint32_t T, float k_near) {
    sphref_params p;
    memset(&p, 0, sizeof p);
    p.particle_radius = 0.05f;
    p.smoothing_radius = 0.35f;
    p.g = 9.8f;
    for (int a = 0; a < 3; ++a) {
        p.min_xyz[a] = mn[a];
        p.max_xyz[a] = mx[a];
    }
    p.dt = 1.0f / 180.0f;
    p.particle_mass = 1.0f;
    p.pressure_multiplier = 500.0f;
    p.target_density = 55.0f;
    p.viscosity_strength = 0.0f;
    const float R = p.smoothing_radius;
    p.volume = (float)(2.0 * 3.14159265359 * pow(R, 5) / 15.0);
    p.near_pressure_multiplier = k_near;
    p.table_cells = T;
    p.substeps = 3;
    return p;
}

static void check(int ok, const char *what) {
    if (!ok) {
        fprintf(stderr, "sanitize: %s failed\n", what);
        exit(1);
    }
}

/* Every FP32 stage by hand, the composed substep with all outputs, the FP64 substep on its
 * decisions, the brute-force density, the colour ramp and the density map. */
static void run(const char *name, const sphref_params *p, uint32_t n, const float lo[3],
                const float hi[3]) {
    const uint32_t T = p->table_cells;
    float *pos = calloc((size_t)4 * n, sizeof(float)), *vel = calloc((size_t)4 * n, sizeof(float));
    float *pp = calloc((size_t)4 * n, sizeof(float)), *pv = calloc((size_t)4 * n, sizeof(float));
    float *ps = calloc((size_t)4 * n, sizeof(float)), *pvs = calloc((size_t)4 * n, sizeof(float));
    float *vs = calloc((size_t)4 * n, sizeof(float)), *vf = calloc((size_t)4 * n, sizeof(float));
    float *dens = calloc(n, sizeof(float)), *near = calloc(n, sizeof(float));
    float *bd = calloc(n, sizeof(float)), *bn = calloc(n, sizeof(float));
    float *col = calloc((size_t)4 * n, sizeof(float));
    uint32_t *keys = calloc(n, sizeof(uint32_t)), *sk = calloc(n, sizeof(uint32_t));
    uint32_t *si = calloc(n, sizeof(uint32_t));
    int32_t *st = calloc(T, sizeof(int32_t)), *en = calloc(T, sizeof(int32_t));
    check(pos && vel && pp && pv && ps && pvs && vs && vf && dens && near && bd && bn && col && keys && sk &&
              si && st && en,
          "allocation");
    fill(n, lo, hi, pos, vel);

    sphref_derived d;
    sphref_derive(p, &d);
    check(d.mask == T - 1u, "derive");

    /* the stages of one substep, as the composed substep runs them */
    sphref_hash(p, n, pos, keys);
    check(sphref_sort(n, T, keys, sk, si) == 0, "sort");
    for (uint32_t j = 1; j < n; ++j) check(sk[j - 1] <= sk[j], "sorted order");
    sphref_table(n, T, sk, st, en);
    for (uint32_t j = 0; j < n; ++j) {
        memcpy(ps + 4 * j, pos + 4 * si[j], 4 * sizeof(float));
        memcpy(vs + 4 * j, vel + 4 * si[j], 4 * sizeof(float));
    }
    sphref_predict(p, n, ps, vs, pp, pv);
    sphref_density(p, n, pp, st, en, dens, near);
    sphref_density_bruteforce(p, n, pp, sk, bd, bn);
    sphref_force(p, n, pp, pv, vs, dens, near, st, en, vf);
    memcpy(pvs, vf, (size_t)16 * n);
    sphref_integrate(p, n, ps, pvs);
    sphref_color(n, pvs, col);
    for (uint32_t j = 0; j < n; ++j) check(isfinite(dens[j]) && isfinite(near[j]), "finite densities");

    /* the composed substep with every output, then the FP64 substep on its decisions */
    float *p1 = malloc((size_t)16 * n), *v1 = malloc((size_t)16 * n);
    float *o_pp = malloc((size_t)16 * n), *o_pv = malloc((size_t)16 * n), *o_vf = malloc((size_t)16 * n);
    float *o_d = malloc((size_t)4 * n), *o_n = malloc((size_t)4 * n);
    uint32_t *o_k = malloc((size_t)4 * n), *o_sk = malloc((size_t)4 * n), *o_si = malloc((size_t)4 * n);
    int32_t *o_st = malloc((size_t)4 * T), *o_en = malloc((size_t)4 * T);
    double *d64 = malloc((size_t)8 * n), *n64 = malloc((size_t)8 * n), *vf64 = malloc((size_t)24 * n);
    double *po64 = malloc((size_t)24 * n), *vo64 = malloc((size_t)24 * n), *sc64 = malloc((size_t)8 * n);
    check(p1 && v1 && o_pp && o_pv && o_vf && o_d && o_n && o_k && o_sk && o_si && o_st && o_en && d64 && n64 &&
              vf64 && po64 && vo64 && sc64,
          "allocation");
    memcpy(p1, pos, (size_t)16 * n);
    memcpy(v1, vel, (size_t)16 * n);
    const sphref_substep_out out = {o_k, o_sk, o_si, o_st, o_en, o_pp, o_pv, o_d, o_n, o_vf};
    check(sphref_substep(p, n, p1, v1, &out) == 0, "substep");
    check(memcmp(o_sk, sk, (size_t)4 * n) == 0 && memcmp(o_d, dens, (size_t)4 * n) == 0, "substep stages");
    check(sphref64_substep(p, n, pos, vel, o_si, o_st, o_en, o_pp, d64, n64, vf64, po64, vo64, sc64) == 0,
          "substep64");

    /* several frames, then the density map of the last substep (voxel not dividing the box) */
    check(sphref_step(p, n, p1, v1, 2) == 0, "step");
    sphref_hash(p, n, p1, keys);
    check(sphref_sort(n, T, keys, sk, si) == 0, "sort after step");
    sphref_table(n, T, sk, st, en);
    for (uint32_t j = 0; j < n; ++j) memcpy(ps + 4 * j, p1 + 4 * si[j], 4 * sizeof(float));
    uint32_t dims[3];
    const float voxel = 0.13f;
    sphref_density_map_dims(p, voxel, dims);
    const size_t voxels = (size_t)dims[0] * dims[1] * dims[2];
    float *map = malloc(voxels * sizeof(float));
    check(map != NULL, "allocation");
    sphref_density_map(p, n, ps, st, en, voxel, map);
    for (size_t v = 0; v < voxels; ++v) check(isfinite(map[v]), "finite density map");
    printf("sanitize: %-28s n=%u T=%u map %ux%ux%u\n", name, n, T, dims[0], dims[1], dims[2]);

    free(map);
    free(p1); free(v1); free(o_pp); free(o_pv); free(o_vf); free(o_d); free(o_n); free(o_k); free(o_sk);
    free(o_si); free(o_st); free(o_en); free(d64); free(n64); free(vf64); free(po64); free(vo64); free(sc64);
    free(pos); free(vel); free(pp); free(pv); free(ps); free(pvs); free(vs); free(vf); free(dens); free(near);
    free(bd); free(bn); free(col); free(keys); free(sk); free(si); free(st); free(en);
}

int main(void) {
    {
        const float mn[3] = {0, 0, 0}, mx[3] = {3, 2, 2}, lo[3] = {0, 0, 0}, hi[3] = {1.5f, 2, 2};
        sphref_params p = scene(mn, mx, 1024, 0.0f);
        run("block, colliding keys", &p, 3000, lo, hi);
    }
    {
        const float mn[3] = {-2.5f, -1.25f, -0.75f}, mx[3] = {0.5f, 1.0f, 1.5f};
        const float lo[3] = {-2.5f, -1.25f, -0.75f}, hi[3] = {-1.0f, 1.0f, 1.5f};
        sphref_params p = scene(mn, mx, 4096, 50.0f);
        run("negative min, near pressure", &p, 2500, lo, hi);
    }
    {
        const float mn[3] = {0, 0, 0}, mx[3] = {1, 1, 1}, lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
        sphref_params p = scene(mn, mx, 1, 50.0f);
        run("one table cell", &p, 400, lo, hi);
    }
    {
        const float mn[3] = {0, 0, 0}, mx[3] = {2, 2, 2}, lo[3] = {1, 1, 1}, hi[3] = {1.01f, 1.01f, 1.01f};
        sphref_params p = scene(mn, mx, 8, 0.0f);
        run("one particle", &p, 1, lo, hi);
    }
    {
        /* a box thinner than the particle radius plus the integrate margin: positions
         * clamp to both walls, cells below zero wrap */
        const float mn[3] = {0, 0, 0}, mx[3] = {0.6f, 0.6f, 0.6f}, lo[3] = {0, 0, 0}, hi[3] = {0.6f, 0.6f, 0.6f};
        sphref_params p = scene(mn, mx, 64, 50.0f);
        run("thin box", &p, 200, lo, hi);
    }
    printf("sanitize: ok\n");
    return 0;
}
