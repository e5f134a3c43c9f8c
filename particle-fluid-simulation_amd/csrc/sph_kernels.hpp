// Host-side launchers of the SPH substep kernels (definitions in sph_kernels.hip and
// radix_sort.hip).  Every launcher enqueues on `stream` and returns hipGetLastError().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.hpp"

namespace sph {

// Device buffers of one substep chain.  Particle arrays are indexed by slot; for a batch,
// scene s owns slots [p_begin, p_end) and table cells [table_base, table_base + T_s).
//
// Two slot orders exist.  Hash order is the reference's: slots sorted by (cell key,
// input index); the table, the sort outputs and the state written back are in it.  The
// per-lane neighbour passes work in cell order instead: whole key runs, each kept in its
// hash order, placed by the grid bucket of the run's first particle, so that particles of
// neighbouring cells sit next to each other in memory and in a wave.  pred_pos,
// pos_sorted, vel_sorted, dens and the debug captures are in cell order on the per-lane
// path and in hash order on the brick path.
struct ChainBuffers {
    float4 *pos;         // state positions (vec4), input order of the substep
    float4 *vel;         // state velocities
    uint32_t *keys;      // cell key per slot (hash pass output)
    uint32_t *sort_k[2]; // radix ping-pong
    uint32_t *sort_v[2];
    float4 *pred_pos;    // predicted positions, sorted order (w: near density once the
                         // list-recording density pass has run, else 1)
    float4 *pos_sorted;  // positions, sorted order (a density map after a per-lane substep
                         // gathers the new ones into it first)
    float4 *vel_sorted;  // velocities, sorted order (w: density once the list-recording
                         // density pass has run)
    float2 *dens;        // (density, near density), sorted order
    int2 *table;         // (start, end) per cell, -1 = empty
    uint32_t *ghist;     // [kGhistCopies][passes][bins] digit histograms (kGhistWords each)
    uint32_t *lookback;  // [passes][tiles][bins] look-back words
    uint32_t *tile_ctr;  // [4] tile counters
    float4 *dbg_pred_vel;   // optional (debug capture): predicted velocities, sorted
    float4 *dbg_vel_force;  // optional (debug capture): velocity after force, sorted
    // Brick work queue (see sph_brick.hip).
    uint32_t *hfull;        // full cell hash of each slot's current position, sorted order
    uint32_t *brick_flag;   // [brick_total] non-empty marks
    uint32_t *queue;        // [brick_total] non-empty bricks of this substep
    uint32_t *stray;        // [n] slots outside every brick grid
    uint32_t *counters;     // [0] queue length [1] strays [2] density drain [3] force drain
                            // [4] entries of the longest force list above list_cap / 2
                            // since the host last read the record (0: none)
                            // [5] tiles streamed in more than one row group [6] tiles handed
                            // over to the strays [7] their particles, listed from the top of
                            // `stray` (the last tiled density pass)
    // Cell-ordered layout (sph_layout.hip).
    uint32_t *bcount;       // [bucket_total] particles per bucket, then the bucket's first slot
    uint32_t *blookback;    // [bucket tiles] look-back words of the bucket scan
    uint2 *run_info;        // [table_total] per occupied key: (bucket, offset in the bucket)
    int2 *crun;             // [table_total] per key: its run's (start, end) in cell order, -1: empty
    uint32_t *hslot;        // [n] hash slot of each cell-ordered slot
    int2 *grid;             // [grid_total] (start, end) in cell order of each halo-grid cell's run
    // Force neighbour lists recorded by the per-lane density pass (cell-ordered slots).
    uint32_t *nlist;        // [list_cap][n]: entry k of slot j at nlist[k * n + j]
    uint8_t *ncount;        // [n] entries, or 0xFF: the force pass walks the cells
    uint32_t list_cap;      // entries per slot allocated at nlist (<= kList)
    // Tiled neighbour passes (sph_tiles.hip): per cell-ordered slot, the tile that owns the
    // particle and the region cell of its home copy, or (kNoTile, 0) for a stray, which is
    // listed at `stray` (count in counters[1]).  Owned particles' force lists hold 16-bit
    // indices into the tile's staged region (entry k of slot j at nlist16()[k * n + j]).
    uint2 *towner;
    __host__ __device__ uint16_t *nlist16() const { return reinterpret_cast<uint16_t *>(nlist); }
};

constexpr uint32_t kNoTile = 0xFFFFFFFFu;

uint32_t radix_tile_size();
// tile_ctr[0..3]: radix passes; [4]: bucket scan.
constexpr uint32_t kTileCounters = 8;
// Radix digit widths: 8, 11 or 12 bits (256, 2048 or 4096 bins), at most 4 passes; the
// digit histograms of all passes, [pass][bins], fit kGhistWords.  The hash pass adds into
// kGhistCopies copies of them (one per XCD, [copy][pass][bins] with a stride of the passes'
// words), which the radix passes sum: each copy takes an eighth of the hashing workgroups'
// additions to one bin, which device-wide atomics would otherwise serialise.
constexpr uint32_t kGhistWords = 3u * 4096u;
constexpr uint32_t kGhistCopies = 8u;

// One stable LSD pass over `bits` (<= digit_bits) bits at `shift`, digit_bits = 8, 11, 12.
hipError_t radix_pass(hipStream_t stream, const uint32_t *kin, const uint32_t *vin,
                      uint32_t *kout, uint32_t *vout, uint32_t n, uint32_t shift,
                      uint32_t bits, const uint32_t *ghist, uint32_t ghist_stride, uint32_t *lookback,
                      uint32_t *tile_counter, uint32_t digit_bits);

// Hash pass: keys of the current positions, digit histograms for `passes` passes, and
// clearing of the cell table, the look-back words and the tile counters; with
// `bucket_total` > 0 also of the bucket counts.
template <class Scenes>
hipError_t launch_hash(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                       uint32_t n, uint32_t table_total, uint32_t passes,
                       uint32_t lookback_words, uint32_t bucket_total, uint32_t digit_bits);

// Cell start/end table from the sorted keys; clears histograms.  With `bricks`, also
// gathers the state into hash order with predicted positions, marks non-empty bricks and
// lists stray particles; without, places every run in its bucket of the cell-ordered
// layout (the state is gathered later by launch_place).
template <class Scenes>
hipError_t launch_reorder(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                          const uint32_t *sorted_keys, const uint32_t *sorted_idx, uint32_t n,
                          bool bricks, uint32_t ghist_words);

// Cell-ordered layout (sph_layout.hip): the bucket scan, then the gather of the state into
// cell order with predicted positions and each key's cell-ordered run, then the halo grids.
constexpr uint32_t kBucketTile = 2048;
hipError_t launch_bucket_scan(hipStream_t stream, const ChainBuffers &b, uint32_t bucket_total);
template <class Scenes>
hipError_t launch_place(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                        const uint32_t *sorted_keys, const uint32_t *sorted_idx, uint32_t n,
                        uint32_t grid_total, bool tiles);

// Launch configurations of the one-lane-per-particle density / force kernels: candidates
// loaded per scan step, workgroup size, and whether the density pass records force
// neighbour lists (worth it unless lists overflow, i.e. in very dense scenes).  The
// autotuner picks one per size bucket.
struct LaneConfig {
    int unroll;
    int block;
    bool lists;
    bool tolerance;  // approximate square root / reciprocal: not bit-exact (opt-in per world)
};
constexpr int kNumLaneConfigs = 12;
constexpr int kNumExactLaneConfigs = 10;  // configurations 0..9 are bit-exact
constexpr int kDefaultLaneConfig = 3;
extern const LaneConfig kLaneConfigs[kNumLaneConfigs];

// Density over the cell-ordered layout.
template <class Scenes>
hipError_t launch_density(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                          uint32_t n, int cfg);

// Pressure + viscosity force, then integrate, over the cell-ordered layout; writes the new
// state into b.pos / b.vel at the hash slots.
template <class Scenes>
hipError_t launch_force(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                        uint32_t n, int cfg);

// Tiled neighbour passes (sph_tiles.hip).  A workgroup stages the cell runs of its tile
// grown by one cell per side in LDS and evaluates the particles the tile owns (predicted
// cell inside the tile, recorded by launch_place); the first `stray_blocks` workgroups walk
// the strays through global memory.  Configurations: workgroup lanes and LDS entries, and
// the force pass: tiled (lists of 16-bit staged indices), or per-lane configuration
// `force_lane` over lists of cell-ordered slots.  The exact configurations come first.
struct TileConfig {
    int block;
    int entries;  // staged entries per tile (a tile with more falls back to global walks)
    bool tolerance;
    int force_lane;  // per-lane force configuration (kLaneConfigs), -1: the tiled force pass
};
constexpr int kNumTileConfigs = 5;
constexpr int kNumExactTileConfigs = 4;
// The configurations the autotuners measure: the tiled force pass (2, 3) lost wherever it
// was measured, and its 16-bit lists hand dense tiles to global walks in both passes, so
// it is only run when pinned.
bool tile_config_tuned(int cfg, bool tolerance);
extern const TileConfig kTileConfigs[kNumTileConfigs];
// Region cells a tile may have ((tw + 2) (th + 2) (td + 2)).
constexpr uint32_t kTileRegionMax = 768;
// Tile dimensions for a scene whose occupied cells hold `ppc` particles on average.
void tile_dims_for(SceneConst &c, double ppc, int cfg);
uint32_t scene_tiles(const SceneConst &c);
template <class Scenes>
hipError_t launch_density_tile(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                               uint32_t tile_total, uint32_t stray_blocks, int cfg);
template <class Scenes>
hipError_t launch_force_tile(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                             uint32_t tile_total, uint32_t stray_blocks, int cfg);

// Brick path (sph_brick.hip): compacts the non-empty bricks into the work queue, then
// persistent workgroups drain it for the density and the force + integrate passes.
// scene_of[j] for slots [0, n) from the scenes' slot ranges (table sorted by p_begin).
hipError_t launch_fill_scene_of(hipStream_t stream, const SceneConst *table, uint32_t count,
                                uint16_t *scene_of, uint32_t n);

void brick_prepare();  // caches occupancy-derived grid sizes; call outside graph capture
hipError_t launch_queue(hipStream_t stream, const ChainBuffers &b, uint32_t brick_total);
template <class Scenes>
hipError_t launch_density_brick(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                                uint32_t n, uint32_t brick_total);
template <class Scenes>
hipError_t launch_force_brick(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                              uint32_t n, uint32_t brick_total);

}  // namespace sph
