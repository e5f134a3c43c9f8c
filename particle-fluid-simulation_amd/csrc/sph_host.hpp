// Host-side pieces shared by the single-world and batched C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_world.h"
#include "sph_kernels.hpp"

namespace sph {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
const char *last_error();

SceneConst make_scene_const(const sph_params &p, uint32_t table_cells, uint32_t table_base,
                            uint32_t p_begin, uint32_t p_end, uint32_t brick_base,
                            uint32_t bucket_base, uint32_t grid_base);
// Bricks of one scene's cell grid (0 when the grid is too large to tile).
uint64_t scene_bricks(const SceneConst &c);
// Buckets of one scene's cell-ordered layout (at most 2n + 1024).
uint32_t scene_buckets(const SceneConst &c);
int validate_params(const sph_params *p);
int validate_state(const sph_params &p, uint32_t n, const float *pos4, const float *vel4);
uint32_t key_bits_for(uint64_t table_total);
void speed_ramp(uint32_t n, const float *vel4, float *color4);

// Force-list entries a chain starts with, the fewest an uploaded state is given and the
// fewest the lists shrink to; records in a row without a list above half the capacity
// before it halves.
constexpr uint32_t kListStart = 64;
constexpr uint32_t kListQuiet = 8;
// Force-list entries for an uploaded state of n particles (vec4 positions).
uint32_t list_entries_for_state(const sph_params &p, uint32_t n, const float *pos4);
// Particles per cell over the bounding box of an uploaded state (sizes the cell tiles).
double particles_per_cell_for_state(const sph_params &p, uint32_t n, const float *pos4);
// Workgroups of a tiled launch that take the strays: `untiled` particles of scenes without
// tiles among n.
uint32_t stray_blocks_for(uint32_t untiled, uint32_t n, int block);

// hipEvents bracketing the five passes of one substep.
struct PassEvents {
    hipEvent_t ev[6] = {};
    int create();
    void destroy();
};

// Buffers and launch sequence of the substep chain over `n` slots and `table_total` cells.
struct Chain {
    ChainBuffers b{};
    uint32_t n = 0, table_total = 0, key_bits = 0, passes = 0, tiles = 0, lookback_words = 0;
    uint32_t brick_total = 0;  // bricks over all scenes; 0 disables the brick kernels
    uint32_t bucket_total = 0; // buckets of the cell-ordered layout over all scenes (>= 1)
    uint32_t grid_total = 0;   // halo-grid cells over all scenes
    bool bricks = false;       // density/force through the brick kernels (else per lane)
    bool cell_order = true;    // the last enqueued substep left its arrays in cell order
    int lane_cfg = kDefaultLaneConfig;  // launch configuration of the per-lane kernels
    int tile_cfg = -1;         // configuration of the tiled kernels, -1: per-lane kernels
    uint32_t tile_total = 0;   // tiles over all scenes (tiled kernels)
    uint32_t stray_blocks = 8; // workgroups of the tiled launches that take the strays
    uint32_t radix_bits = 8;   // radix digit width (8, 11 or 12)
    uint64_t lookback_cap = 0; // look-back words allocated (fixes which widths are possible)
    uint32_t n_cap = 0, t_cap = 0, b_cap = 0, u_cap = 0, g_cap = 0;  // allocated capacity
    uint64_t bytes = 0;        // device memory allocated by allocate() and set_list_cap()
    uint32_t *h_list = nullptr;  // pinned host words: [0] the step calls' copy of the list record, [1] a sync's

    // Allocates for the given capacity (force lists of list_entries per slot) and
    // configures the chain to use all of it.
    int allocate(uint32_t n, uint32_t table_total, uint32_t brick_total, uint32_t bucket_total,
                 uint32_t grid_total, uint32_t list_entries);
    // Reallocates the force lists at `entries` per slot (1..kList) for n_cap slots: the new
    // buffer is allocated before the old one is freed, so on failure the chain keeps the
    // old capacity.  The caller has synchronised the chain's stream.
    int set_list_cap(uint32_t entries);
    // Force-list sizing from the record the density passes leave (the longest list above
    // half the capacity since the record was last read and cleared): `list_target` gives the
    // capacity a record asks for (0: keep), twice
    // the capacity or the list plus a quarter (in steps of 32, at most kList) past a list that
    // did not fit, and, with `shrink`, half the capacity (at least kListStart) after
    // kListQuiet records in a row with no list above half of it.
    uint32_t list_target(uint32_t record, bool shrink);
    // Synchronises `s`; with `adapt`, first queues the record's copy and afterwards resizes
    // the lists as it asks.  Returns 1 when the capacity changed (a captured graph must be
    // dropped), 0, or an error.
    int sync_lists(hipStream_t s, bool adapt);
    // Between the frames of a step call, without waiting: looks at the record the copy
    // queued at the previous call left (if it has landed) and queues the next copy; when the
    // record asks for longer lists it synchronises and grows them (shrinking waits for
    // sync_lists).  Returns 1 when the capacity changed, 0, or an error.
    int poll_lists_async(hipStream_t s);
    uint32_t list_quiet = 0;        // records in a row with no list above half the capacity
    hipEvent_t list_ev = nullptr;   // marks the landing of the queued record copy
    bool list_poll_pending = false;
    // Uses the first n slots, table_total cells, brick_total bricks, bucket_total buckets
    // and grid_total halo-grid cells (within capacity).
    void configure(uint32_t n, uint32_t table_total, uint32_t brick_total, uint32_t bucket_total,
                   uint32_t grid_total);
    // Look-back words a radix width needs at the configured size.
    uint64_t lookback_words_for(uint32_t digit_bits) const;
    // Selects the radix digit width; false when it is not 8, 11 or 12 or its look-back
    // words exceed what was allocated.
    bool set_radix_bits(uint32_t digit_bits);
    uint32_t ghist_words() const { return passes << radix_bits; }
    int set_debug(bool on);
    void release();
    const uint32_t *sorted_keys() const;
    const uint32_t *sorted_idx() const;  // nullptr means identity (no sort pass ran)

    // The neighbour passes run the tiled kernels.
    bool tiled() const { return !bricks && tile_cfg >= 0 && tile_total > 0; }

    template <class Scenes>
    hipError_t enqueue(hipStream_t s, const Scenes &sc, PassEvents *ev);
};

// One instantiated graph of the substep chain; dropped whenever what it captured changes.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int ensure_begin(hipStream_t s);
    int end(hipStream_t s, hipError_t launch_status);
    void drop();
};

// Copies one intermediate buffer of the last substep into `dst` in the reference's format,
// restricted to slots [p_begin, p_end) and table cells [t_begin, t_end); slot indices and
// keys are made scene-local.
int dump_stage(const Chain &c, hipStream_t s, int stage, uint32_t p_begin, uint32_t p_end,
               uint32_t t_begin, uint32_t t_end, void *dst, size_t bytes);

}  // namespace sph
