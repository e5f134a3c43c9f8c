/*
 * CPU oracle for the per-substep SPH chain of mojofill/particle-fluid-simulation.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or calls this
 * code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
 * only as the checker (or as the timed CPU baseline), never as the thing measured or
 * shipped.
 *
 * Two restatements of the reference's compute shaders are provided:
 *   - FP32 (sph_ref32.c): every float operation in the shaders' source order, each one
 *     rounded to binary32 (built with -ffp-contract=off, no fast-math), IEEE-correct
 *     division and sqrt.  Integer stages (cell keys, stable sort, start/end table) are
 *     exact u32/i32 arithmetic.
 *   - FP64 (sph_ref64.c): the same equations in binary64.  It takes every integer decision
 *     (sorted order, start/end table, the cell keys each particle visits) from an FP32 run,
 *     so its difference to the FP32 run measures precision only.
 *
 * Reference files restated (paths relative to the reference repository root):
 *   src/shaders/predictPositions.comp, updateCellEntries.comp,
 *   multiRadixSortHistograms.comp + multiRadixSort.comp (semantics: stable LSD sort),
 *   scatterPass.comp, clearStartIndices.comp, updateStartIndices.comp,
 *   calculateDensities.comp, applyPressureAndViscosityForce.comp, updateVelocities.comp,
 *   updateDensityMap.comp (the density map after the substeps),
 *   src/renderer/commands.c:106-161,193-402 (pass order), src/main.c:37-82,349 (params).
 *
 * Pins (tests/test_oracle_*.py): closed-form hash known answers, brute-force O(N^2)
 * neighbour sums with the reference's cell-visit rules, isolated-particle and two-particle
 * closed forms, sort invariants against Python's stable sort, per-axis predict/integrate
 * closed forms, momentum antisymmetry of the pressure term.  Parity of the last-bit choices
 * lavapipe would make (FMA contraction in dot/distance, pow, vector/scalar division) is
 * unpinned: no run of the reference is possible here; the tolerance absorbs them.
 *
 * Neighbour-velocity semantics: the reference's force shader reads velocities[i] of
 * neighbours while other invocations add to velocities[j] in the same dispatch
 * (applyPressureAndViscosityForce.comp:313 vs :336), a read/write race.  This oracle uses
 * Jacobi semantics: every neighbour read sees the pre-force velocity.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field meaning and order as sph_params in include/sph_world.h, i.e. the physics
 * fields of UniformBufferObject (src/renderer/vk_types.h:52-65) plus the table size and
 * substep count. */
typedef struct sphref_params {
    float particle_radius;
    float smoothing_radius;
    float g;
    float min_xyz[3];
    float max_xyz[3];
    float dt;
    float particle_mass;
    float pressure_multiplier;
    float target_density;
    float viscosity_strength;
    float volume;
    float near_pressure_multiplier;
    uint32_t table_cells;
    uint32_t substeps;
} sphref_params;

/* Constants the shaders derive from the UBO, computed once on the host in float. */
typedef struct sphref_derived {
    uint32_t mask;           /* table_cells - 1                                        */
    float R2;                /* R*R                       (applyPressure...comp:282)    */
    float slope_scale;       /* 15/(PI*pow(R,5))          (applyPressure...comp:243)    */
    float near_slope_scale;  /* -45/(PI*pow(R,6))         (applyPressure...comp:258)    */
    float near_vol_inv;      /* 15/(PI*pow(R,6))          (calculateDensities.comp:110) */
    float integrate_margin;  /* particleRadius + 0.5      (updateVelocities.comp:81)    */
} sphref_derived;

/* Everything one substep produces, in the reference's buffer formats.  Arrays are
 * caller-owned; any pointer may be NULL to skip that output. */
typedef struct sphref_substep_out {
    uint32_t *cell_keys;     /* [n] key per particle in input order (CellEntry.key)     */
    uint32_t *sorted_keys;   /* [n] sorted CellEntry.key                               */
    uint32_t *sorted_idx;    /* [n] sorted CellEntry.particleIndex (input index)       */
    int32_t  *start;         /* [T] startIndices, -1 = empty                           */
    int32_t  *end;           /* [T] endIndices,   -1 = empty                           */
    float    *pred_pos;      /* [4n] predictedPositions, sorted order                  */
    float    *pred_vel;      /* [4n] predictedVelocities, sorted order                 */
    float    *density;       /* [n] densities, sorted order                            */
    float    *near_density;  /* [n] nearDensities, sorted order                        */
    float    *vel_force;     /* [4n] velocities after the force pass, sorted order     */
} sphref_substep_out;

void sphref_derive(const sphref_params *p, sphref_derived *d);

uint32_t sphref_cell_hash(uint32_t x, uint32_t y, uint32_t z);

/* --- FP32 restatement (sph_ref32.c) --------------------------------------------------- */
void sphref_predict(const sphref_params *p, uint32_t n, const float *pos4, const float *vel4,
                    float *pred_pos4, float *pred_vel4);
void sphref_hash(const sphref_params *p, uint32_t n, const float *pos4, uint32_t *keys);
/* Stable sort of (key, i) pairs by key; keys < table_cells.  out_idx[j] = input index. */
int  sphref_sort(uint32_t n, uint32_t table_cells, const uint32_t *keys, uint32_t *out_keys,
                 uint32_t *out_idx);
void sphref_table(uint32_t n, uint32_t table_cells, const uint32_t *sorted_keys,
                  int32_t *start, int32_t *end);
void sphref_density(const sphref_params *p, uint32_t n, const float *pred_pos4,
                    const int32_t *start, const int32_t *end, float *density,
                    float *near_density);
void sphref_force(const sphref_params *p, uint32_t n, const float *pred_pos4,
                  const float *pred_vel4, const float *vel4, const float *density,
                  const float *near_density, const int32_t *start, const int32_t *end,
                  float *vel_out4);
void sphref_integrate(const sphref_params *p, uint32_t n, float *pos4, float *vel4);
void sphref_color(uint32_t n, const float *vel4, float *color4);

/* updateDensityMap.comp:114-148: density / volume on voxels of edge voxel_size over the
 * box, x fastest, from positions in sorted order and the start/end table of the substep
 * that produced them.  dims: the voxel counts (renderer.c:774-783). */
void sphref_density_map_dims(const sphref_params *p, float voxel_size, uint32_t dims[3]);
void sphref_density_map(const sphref_params *p, uint32_t n, const float *pos4,
                        const int32_t *start, const int32_t *end, float voxel_size,
                        float *out);

/* One substep in place on pos4/vel4 (n vec4 each); the arrays end up in sorted order.
 * Returns 0, or -1 on allocation failure / invalid sizes. */
int sphref_substep(const sphref_params *p, uint32_t n, float *pos4, float *vel4,
                   const sphref_substep_out *out);
/* frames * substeps substeps. */
int sphref_step(const sphref_params *p, uint32_t n, float *pos4, float *vel4, uint32_t frames);

/* Brute-force neighbour sums (O(n^2)) that follow the same cell-visit rules, for pinning. */
void sphref_density_bruteforce(const sphref_params *p, uint32_t n, const float *pred_pos4,
                               const uint32_t *sorted_keys, float *density,
                               float *near_density);

/* --- FP64 oracle (sph_ref64.c) -------------------------------------------------------- */
/* Given the FP32 run's integer decisions (sorted order, table, and the FP32 predicted
 * positions used to choose visited cells), recompute one substep in binary64.
 * pos4/vel4 are the substep's input state (input order); outputs are in sorted order.
 * dv_scale (may be NULL) receives, per particle, the sum over its force pairs of the
 * magnitudes of the pair's pressure and viscosity contributions to the velocity change,
 * (|a + c| + |b| |v_i - u|) dt / rho: the scale against which the error of the net change
 * (where pair terms cancel) is measured. */
int sphref64_substep(const sphref_params *p, uint32_t n, const float *pos4, const float *vel4,
                     const uint32_t *sorted_idx, const int32_t *start, const int32_t *end,
                     const float *pred_pos4_f32, double *density, double *near_density,
                     double *vel_force3, double *pos_out3, double *vel_out3, double *dv_scale);

#ifdef __cplusplus
}
#endif
#endif
