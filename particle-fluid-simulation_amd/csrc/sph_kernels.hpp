// Host-side launchers of the SPH substep kernels (definitions in sph_kernels.hip and
// radix_sort.hip).  Every launcher enqueues on `stream` and returns hipGetLastError().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.hpp"

namespace sph {

// Device buffers of one substep chain.  Particle arrays are indexed by slot; for a batch,
// scene s owns slots [p_begin, p_end) and table cells [table_base, table_base + T_s).
struct ChainBuffers {
    float4 *pos;         // state positions (vec4), input order of the substep
    float4 *vel;         // state velocities
    uint32_t *keys;      // cell key per slot (hash pass output)
    uint32_t *sort_k[2]; // radix ping-pong
    uint32_t *sort_v[2];
    float4 *pred_pos;    // predicted positions, sorted order
    float4 *pos_sorted;  // positions, sorted order
    float4 *vel_sorted;  // velocities, sorted order
    float2 *dens;        // (density, near density), sorted order
    int2 *table;         // (start, end) per cell, -1 = empty
    uint32_t *ghist;     // [4][256] digit histograms
    uint32_t *lookback;  // [4][tiles][256] look-back words
    uint32_t *tile_ctr;  // [4] tile counters
    float4 *dbg_pred_vel;   // optional (debug capture): predicted velocities, sorted
    float4 *dbg_vel_force;  // optional (debug capture): velocity after force, sorted
};

uint32_t radix_tile_size();

hipError_t radix_pass(hipStream_t stream, const uint32_t *kin, const uint32_t *vin,
                      uint32_t *kout, uint32_t *vout, uint32_t n, uint32_t shift,
                      uint32_t bits, const uint32_t *ghist, uint32_t *lookback,
                      uint32_t *tile_counter);

// Hash pass: keys of the current positions, digit histograms for `passes` passes, and
// clearing of the cell table, the look-back words and the tile counters.
template <class Scenes>
hipError_t launch_hash(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                       uint32_t n, uint32_t table_total, uint32_t passes,
                       uint32_t lookback_words);

// Gather into sorted order, predicted positions, cell start/end table; clears histograms.
template <class Scenes>
hipError_t launch_reorder(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                          const uint32_t *sorted_keys, const uint32_t *sorted_idx, uint32_t n);

template <class Scenes>
hipError_t launch_density(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                          uint32_t n);

// Pressure + viscosity force, then integrate; writes the new state into b.pos / b.vel.
template <class Scenes>
hipError_t launch_force(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                        uint32_t n);

}  // namespace sph
