// Device-side definitions shared by the SPH kernels (gfx950, wave64).
//
// All float arithmetic in this library is compiled with -ffp-contract=off and IEEE
// division / sqrt, and every expression keeps the operand order of the reference's GLSL,
// so results are reproducible bit for bit and independent of how particles are packed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// Per-scene constants.  One copy per scene: a kernel argument for a single world, an
// array indexed by scene id for a batch.
struct SceneConst {
    float minx, miny, minz;
    float maxx, maxy, maxz;
    float R;                 // smoothing radius
    float R2;                // R*R (force cut-off, applyPressure...comp:282)
    float R2_pre;            // d2 > R2_pre implies sqrt(d2) >= R (density pre-filter)
    float g, dt, m;
    float rho0, k, kn;       // target density, pressure and near-pressure multipliers
    float volume;            // 2*pi*R^5/15
    float slope_scale;       // 15/(PI*pow(R,5))
    float near_slope_scale;  // -45/(PI*pow(R,6))
    float near_vol_inv;      // 15/(PI*pow(R,6))
    float margin_predict;    // particleRadius
    float margin_integrate;  // particleRadius + 0.5
    uint32_t mask;           // table_cells - 1
    uint32_t table_base;     // first cell of this scene in the table
    uint32_t p_begin;        // first particle of this scene
    uint32_t p_end;          // one past the last particle
    int32_t gx, gy, gz;      // cells per axis inside the box: floor((max-min)/R) + 1
    uint32_t nbx, nby, nbz;  // bricks per axis (brick = kBrick^3 cells)
    uint32_t brick_base;     // first global brick id of this scene
    uint32_t use_bricks;     // 1: brick kernels; 0: one lane per particle (sparse scenes)
    // Buckets of the cell-ordered layout: cells grouped 2^bshift per axis, numbered
    // x-fastest; this scene owns buckets [bucket_base, bucket_base + bgx*bgy*bgz).
    uint32_t bucket_base;
    uint32_t bshift;
    uint32_t bgx, bgy;
    // Cell-ordered run of every cell of the grid grown by one halo cell per side,
    // (gx+2) x (gy+2) x (gz+2) entries from grid_base; grid_cells = 0: no such grid (the
    // box is large for the particle count) and every lookup goes through the hash table.
    uint32_t grid_base;
    uint32_t grid_cells;
    // Cell tiles of the tiled neighbour passes: tiles of tw x th x td cells cover the box
    // grid, ntx x nty x ntz of them, numbered x-fastest from tile_base.  use_tiles = 0: the
    // scene's particles take the per-lane code inside the tiled launches.
    uint32_t tile_base;
    uint32_t tw, th, td;
    uint32_t ntx, nty, ntz;
    uint32_t use_tiles;
    uint32_t tile_emax;  // staged entries beyond which a tile's region streams in row groups
};

// cellCoordToHash (updateCellEntries.comp:36-52): u32 arithmetic, wraps mod 2^32.
__device__ __forceinline__ uint32_t cell_hash(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t h = (x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u);
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// floor((v - lo) / R) as a cell coordinate; true division, as the reference divides.
__device__ __forceinline__ int32_t cell_of(float v, float lo, float R) {
    return (int32_t)floorf((v - lo) / R);
}

// Bucket of a run whose particles sit in cell (cx, cy, cz): the cell is clamped to the
// scene's grid (positions clamped outside the box by the walls still get a bucket).
__device__ __forceinline__ uint32_t bucket_of(const SceneConst &c, int32_t cx, int32_t cy, int32_t cz) {
    const uint32_t x = (uint32_t)min(max(cx, 0), c.gx - 1) >> c.bshift;
    const uint32_t y = (uint32_t)min(max(cy, 0), c.gy - 1) >> c.bshift;
    const uint32_t z = (uint32_t)min(max(cz, 0), c.gz - 1) >> c.bshift;
    return c.bucket_base + x + c.bgx * (y + c.bgy * z);
}

// Wall clamp of predictPositions.comp:81-113 / updateVelocities.comp:83-115.
// `vel *= 1.00` in the reference is exact and omitted.
__device__ __forceinline__ void wall_clamp(const SceneConst &c, float margin, float3 &p, float3 &v) {
    if (p.x < c.minx + margin) { p.x = c.minx + margin; v.x = v.x * -1.0f; }
    else if (p.x > c.maxx - margin) { p.x = c.maxx - margin; v.x = v.x * -1.0f; }
    if (p.y < c.miny + margin) { p.y = c.miny + margin; v.y = v.y * -1.0f; }
    else if (p.y > c.maxy - margin) { p.y = c.maxy - margin; v.y = v.y * -1.0f; }
    if (p.z < c.minz + margin) { p.z = c.minz + margin; v.z = v.z * -1.0f; v.z = v.z * 0.95f; }
    else if (p.z > c.maxz - margin) { p.z = c.maxz - margin; v.z = v.z * -1.0f; }
}

// Gravity kick + drift + clamp: predictPositions.comp:74-113 (margin r, dt_scale absent)
// and updateVelocities.comp:75-115 (margin r + 0.5, dt_scale = 1.0).
template <bool kIntegrate>
__device__ __forceinline__ void kick_drift(const SceneConst &c, float3 &p, float3 &v) {
    const float dt = c.dt;
    v.x = v.x + 0.0f * dt;
    v.y = v.y + 0.0f * dt;
    v.z = v.z + (-c.g) * dt;
    if (kIntegrate) {
        p.x = p.x + v.x * dt * 1.0f;
        p.y = p.y + v.y * dt * 1.0f;
        p.z = p.z + v.z * dt * 1.0f;
        wall_clamp(c, c.margin_integrate, p, v);
    } else {
        p.x = p.x + v.x * dt;
        p.y = p.y + v.y * dt;
        p.z = p.z + v.z * dt;
        wall_clamp(c, c.margin_predict, p, v);
    }
}

// Scene lookup policies.  A single world passes its constants by value; a batch reads
// them from a device table through the particle's scene id.
struct OneScene {
    SceneConst c;
    __device__ __forceinline__ const SceneConst &scene(uint32_t) const { return c; }
    __device__ __forceinline__ const SceneConst &scene_of_brick(uint32_t) const { return c; }
    __device__ __forceinline__ const SceneConst &scene_of_grid(uint32_t) const { return c; }
    __device__ __forceinline__ const SceneConst &scene_of_grid_near(uint32_t, uint32_t) const { return c; }
    __device__ __forceinline__ const SceneConst &scene_of_tile(uint32_t) const { return c; }
};

struct ManyScenes {
    const SceneConst *__restrict__ table;
    const uint16_t *__restrict__ scene_of;  // scene id per particle slot
    uint32_t count;                         // number of scenes
    __device__ __forceinline__ SceneConst scene(uint32_t i) const { return table[scene_of[i]]; }
    // Scene whose brick range [brick_base, next brick_base) holds global brick id `b`.
    __device__ __forceinline__ SceneConst scene_of_brick(uint32_t b) const {
        uint32_t lo = 0, hi = count;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (table[mid].brick_base <= b) lo = mid; else hi = mid;
        }
        return table[lo];
    }
    // Scene whose grid range [grid_base, grid_base + grid_cells) holds grid entry `g`
    // (scenes without a grid share the next scene's grid_base).
    __device__ __forceinline__ SceneConst scene_of_grid(uint32_t g) const {
        uint32_t lo = 0, hi = count;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (table[mid].grid_base <= g) lo = mid; else hi = mid;
        }
        return table[lo];
    }
    // Scene whose tile range [tile_base, next tile_base) holds global tile id `t`.
    __device__ __forceinline__ SceneConst scene_of_tile(uint32_t t) const {
        uint32_t lo = 0, hi = count;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (table[mid].tile_base <= t) lo = mid; else hi = mid;
        }
        return table[lo];
    }
    // The same for a grid entry `g` at or after `g0`, an entry the whole wave shares: the
    // search runs once per wave on g0 (scalar loads), and a lane steps forward only across
    // the scene boundaries between g0 and g.
    __device__ __forceinline__ SceneConst scene_of_grid_near(uint32_t g0, uint32_t g) const {
        uint32_t lo = 0, hi = count;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (table[mid].grid_base <= g0) lo = mid; else hi = mid;
        }
        while (lo + 1u < count && table[lo + 1u].grid_base <= g) ++lo;
        return table[lo];
    }
};

// Edge of a brick in cells.  A workgroup stages the (kBrick+2)^3 halo cells of its brick
// in LDS and evaluates every particle whose current cell lies in the brick.
constexpr int kBrick = 4;

// Force neighbour lists: entries per particle (a count fits one byte, 0xFF marks "no
// list"), so that the densest packed scenes (a few hundred candidates, up to ~200
// neighbours) still read a list rather than walk their 27 cells.
constexpr int kList = 254;
constexpr uint8_t kListInvalid = 0xFF;

// The neighbour walk of the density pass (calculateDensities.comp:114-131) for sample
// point s.  Per axis only three sample coordinates exist (s + R*{-1,0,1}), so their cell
// coordinates and in-box flags are computed once; the 27 cells are then visited in the
// reference's (dz, dy, dx) order.  `visit(cx, cy, cz)` is called for each non-skipped cell.
struct AxisCells {
    uint32_t mid;     // cell of s
    uint32_t lo_off;  // cell(s - R) - mid (mod 2^32)
    uint32_t hi_off;  // cell(s + R) - mid (mod 2^32)
    uint32_t inmask;  // bit d set: sample coordinate d lies inside [lo, hi]
    // Arithmetic selection (no per-field selects) keeps the struct in registers when d is
    // not a compile-time constant.
    __device__ __forceinline__ int32_t c(int d) const {
        return (int32_t)(mid + (uint32_t)(d == 0) * lo_off + (uint32_t)(d == 2) * hi_off);
    }
    __device__ __forceinline__ bool in(int d) const { return (inmask >> d) & 1u; }
};

__device__ __forceinline__ AxisCells sample_axis(float s, float lo, float hi, float R) {
    AxisCells a;
    const float q0 = s + R * -1.0f, q1 = s + R * 0.0f, q2 = s + R * 1.0f;
    a.inmask = ((!(q0 < lo) && !(q0 > hi)) ? 1u : 0u) | ((!(q1 < lo) && !(q1 > hi)) ? 2u : 0u) |
               ((!(q2 < lo) && !(q2 > hi)) ? 4u : 0u);
    a.mid = (uint32_t)cell_of(q1, lo, R);
    a.lo_off = (uint32_t)cell_of(q0, lo, R) - a.mid;
    a.hi_off = (uint32_t)cell_of(q2, lo, R) - a.mid;
    return a;
}

}  // namespace sph
