// C-ABI of the single-scene SPH world (include/sph_world.h) and the host orchestration of
// one substep: hash -> radix passes -> reorder/table -> density -> force/integrate, captured
// once into a hipGraph and replayed for every substep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sph_world.h"
#include "sph_map.hpp"
#include "sph_host.hpp"
#include "sph_kernels.hpp"
#include "sph_pairs.hpp"

namespace sph {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(SPH_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

const char *last_error() { return g_last_error.c_str(); }

// Constants derived from the parameter block, computed on the host in binary32 exactly as
// the shaders derive them (calculateDensities.comp:110, applyPressure...comp:243,258,282).
SceneConst make_scene_const(const sph_params &p, uint32_t table_cells, uint32_t table_base,
                            uint32_t p_begin, uint32_t p_end, uint32_t brick_base,
                            uint32_t bucket_base, uint32_t grid_base) {
    const float kPi = 3.14159265359f;
    SceneConst c{};
    const float R = p.smoothing_radius;
    c.minx = p.min_xyz[0]; c.miny = p.min_xyz[1]; c.minz = p.min_xyz[2];
    c.maxx = p.max_xyz[0]; c.maxy = p.max_xyz[1]; c.maxz = p.max_xyz[2];
    c.R = R;
    c.R2 = R * R;
    // Smallest float above R^2 (1 + 2^-20): any d2 beyond it has a correctly rounded
    // sqrt >= R, so skipping those candidates before the sqrt changes no result.
    const double r2 = (double)R * (double)R * (1.0 + std::ldexp(1.0, -20));
    c.R2_pre = std::nextafter((float)r2, INFINITY);
    c.g = p.g;
    c.dt = p.dt;
    c.m = p.particle_mass;
    c.rho0 = p.target_density;
    c.k = p.pressure_multiplier;
    c.kn = p.near_pressure_multiplier;
    c.volume = p.volume;
    c.slope_scale = 15.0f / (kPi * powf(R, 5.0f));
    c.near_slope_scale = -45.0f / (kPi * powf(R, 6.0f));
    c.near_vol_inv = 15.0f / (kPi * powf(R, 6.0f));
    c.margin_predict = p.particle_radius;
    c.margin_integrate = p.particle_radius + 0.5f;
    c.mask = table_cells - 1u;
    c.table_base = table_base;
    c.p_begin = p_begin;
    c.p_end = p_end;
    // Cells a position inside the box can fall in: floor((max - min) / R) + 1 per axis,
    // evaluated in float exactly as the kernels evaluate a particle's cell.
    c.gx = (int32_t)floorf((c.maxx - c.minx) / R) + 1;
    c.gy = (int32_t)floorf((c.maxy - c.miny) / R) + 1;
    c.gz = (int32_t)floorf((c.maxz - c.minz) / R) + 1;
    c.nbx = (uint32_t)((c.gx + kBrick - 1) / kBrick);
    c.nby = (uint32_t)((c.gy + kBrick - 1) / kBrick);
    c.nbz = (uint32_t)((c.gz + kBrick - 1) / kBrick);
    c.brick_base = brick_base;
    c.use_bricks = 1u;
    // Buckets of the cell-ordered layout: single cells unless the grid has more than
    // 2n + 1024 of them, then blocks of 2^s cells per axis until it has no more.
    const double cap = 2.0 * (double)(p_end - p_begin) + 1024.0;
    uint32_t sh = 0;
    for (;; ++sh) {
        const double bx = (double)(((uint32_t)c.gx - 1u) >> sh) + 1.0;
        const double by = (double)(((uint32_t)c.gy - 1u) >> sh) + 1.0;
        const double bz = (double)(((uint32_t)c.gz - 1u) >> sh) + 1.0;
        if (bx * by * bz <= cap || sh == 31u) break;
    }
    c.bshift = sh;
    c.bgx = (((uint32_t)c.gx - 1u) >> sh) + 1u;
    c.bgy = (((uint32_t)c.gy - 1u) >> sh) + 1u;
    c.bucket_base = bucket_base;
    // Halo grid when it has at most 16n + 4096 cells (filling a cell costs about what a
    // particle's 27 hashed lookups would, so a grid pays up to ~20 cells per particle).
    const double halo = ((double)c.gx + 2.0) * ((double)c.gy + 2.0) * ((double)c.gz + 2.0);
    c.grid_base = grid_base;
    c.grid_cells = halo <= 16.0 * (double)(p_end - p_begin) + 4096.0 ? (uint32_t)halo : 0u;
    return c;
}

uint64_t scene_bricks(const SceneConst &c) {
    const uint64_t b = (uint64_t)c.nbx * c.nby * c.nbz;
    return b > (1ull << 28) ? 0 : b;
}

uint32_t scene_buckets(const SceneConst &c) {
    return c.bgx * c.bgy * ((((uint32_t)c.gz - 1u) >> c.bshift) + 1u);
}

int validate_params(const sph_params *p) {
    if (!p) return fail(SPH_ERR_INVALID_ARG, "params is NULL");
    const float vals[] = {p->particle_radius, p->smoothing_radius, p->g, p->dt,
                          p->particle_mass, p->pressure_multiplier, p->target_density,
                          p->volume, p->near_pressure_multiplier,
                          p->min_xyz[0], p->min_xyz[1], p->min_xyz[2],
                          p->max_xyz[0], p->max_xyz[1], p->max_xyz[2]};
    for (float v : vals)
        if (!std::isfinite(v)) return fail(SPH_ERR_INVALID_ARG, "non-finite parameter");
    if (!(p->smoothing_radius > 0.0f)) return fail(SPH_ERR_INVALID_ARG, "smoothing_radius must be > 0");
    if (!(p->volume > 0.0f)) return fail(SPH_ERR_INVALID_ARG, "volume must be > 0");
    for (int a = 0; a < 3; ++a)
        if (!(p->max_xyz[a] > p->min_xyz[a])) return fail(SPH_ERR_INVALID_ARG, "empty box");
    // Cell coordinates must fit in int32 (floor((max-min)/R) is converted to int).
    for (int a = 0; a < 3; ++a)
        if ((double)(p->max_xyz[a] - p->min_xyz[a]) / p->smoothing_radius > 1.0e9)
            return fail(SPH_ERR_INVALID_ARG, "box too large for 32-bit cell coordinates");
    if (p->substeps == 0) return fail(SPH_ERR_INVALID_ARG, "substeps must be >= 1");
    return SPH_OK;
}

int validate_state(const sph_params &p, uint32_t n, const float *pos4, const float *vel4) {
    if (!pos4 || !vel4) return fail(SPH_ERR_INVALID_ARG, "state pointer is NULL");
    for (uint32_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            const float x = pos4[4 * (size_t)i + a], v = vel4[4 * (size_t)i + a];
            if (!std::isfinite(x) || !std::isfinite(v))
                return fail(SPH_ERR_INVALID_ARG, "particle %u: non-finite position or velocity", i);
            if (x < p.min_xyz[a] || x > p.max_xyz[a])
                return fail(SPH_ERR_INVALID_ARG, "particle %u: position outside the box", i);
        }
    }
    return SPH_OK;
}

uint32_t key_bits_for(uint64_t table_total) {
    uint32_t b = 0;
    while ((1ull << b) < table_total) ++b;
    return b;
}

// ---------------------------------------------------------------------------------------
// Chain: buffers and launches of one substep, shared by the single world and the batch.

void Chain::configure(uint32_t n_, uint32_t table_total_, uint32_t brick_total_,
                      uint32_t bucket_total_, uint32_t grid_total_) {
    n = n_;
    table_total = table_total_;
    brick_total = brick_total_;
    bucket_total = bucket_total_;
    grid_total = grid_total_;
    key_bits = key_bits_for(table_total);
    passes = (key_bits + radix_bits - 1u) / radix_bits;
    tiles = (n + radix_tile_size() - 1u) / radix_tile_size();
    lookback_words = passes * tiles << radix_bits;
}

uint64_t Chain::lookback_words_for(uint32_t w) const {
    return (uint64_t)((key_bits + w - 1u) / w) * tiles << w;
}

bool Chain::set_radix_bits(uint32_t w) {
    if (w != 8u && w != 11u && w != 12u) return false;
    if (lookback_words_for(w) > lookback_cap) return false;
    radix_bits = w;
    passes = (key_bits + radix_bits - 1u) / radix_bits;
    lookback_words = passes * tiles << radix_bits;
    return true;
}

int Chain::allocate(uint32_t n_, uint32_t table_total_, uint32_t brick_total_,
                    uint32_t bucket_total_, uint32_t grid_total_, uint32_t list_entries) {
    n_cap = n_;
    t_cap = table_total_;
    b_cap = brick_total_;
    u_cap = bucket_total_;
    g_cap = grid_total_;
    radix_bits = 8;
    configure(n_, table_total_, brick_total_, bucket_total_, grid_total_);
    bricks = brick_total > 0;
    // look-back words for every radix width whose words stay within 16 B per particle (or
    // 8 MiB): wider digits mean fewer passes but 2^w words per tile
    lookback_cap = lookback_words_for(8);
    for (uint32_t w : {11u, 12u}) {
        const uint64_t words = lookback_words_for(w);
        if (words * 4 <= std::max<uint64_t>(16ull * n, 8ull << 20) && words > lookback_cap) lookback_cap = words;
    }
    struct Req { void **p; size_t bytes; };
    const size_t N = n;
    Req reqs[] = {
        {(void **)&b.pos, N * 16}, {(void **)&b.vel, N * 16}, {(void **)&b.keys, N * 4},
        {(void **)&b.sort_k[0], N * 4}, {(void **)&b.sort_v[0], N * 4},
        {(void **)&b.sort_k[1], N * 4}, {(void **)&b.sort_v[1], N * 4},
        {(void **)&b.pred_pos, N * 16}, {(void **)&b.pos_sorted, N * 16},
        {(void **)&b.vel_sorted, N * 16}, {(void **)&b.dens, N * 8},
        {(void **)&b.table, (size_t)table_total * 8}, {(void **)&b.ghist, kGhistCopies * kGhistWords * 4},
        {(void **)&b.lookback, (size_t)(lookback_cap ? lookback_cap : 1) * 4},
        {(void **)&b.tile_ctr, kTileCounters * 4}, {(void **)&b.hfull, N * 4}, {(void **)&b.stray, N * 4},
        {(void **)&b.bcount, (size_t)bucket_total * 4},
        {(void **)&b.blookback, (size_t)((bucket_total + kBucketTile - 1u) / kBucketTile) * 4},
        {(void **)&b.run_info, (size_t)table_total * 8}, {(void **)&b.crun, (size_t)table_total * 8},
        {(void **)&b.hslot, N * 4},
        {(void **)&b.grid, (size_t)(grid_total ? grid_total : 1) * 8},
        {(void **)&b.nlist, N * list_entries * 4}, {(void **)&b.ncount, N}, {(void **)&b.towner, N * 8},
        {(void **)&b.counters, 8 * 4},
        {(void **)&b.brick_flag, (size_t)(brick_total ? brick_total : 1) * 4},
        {(void **)&b.queue, (size_t)(brick_total ? brick_total : 1) * 4},
    };
    if (list_entries < 1u || list_entries > (uint32_t)kList)
        return fail(SPH_ERR_INVALID_ARG, "force lists of %u entries (1..%d)", list_entries, kList);
    bytes = 0;
    for (auto &r : reqs) {
        hipError_t e = hipMalloc(r.p, r.bytes);
        if (e != hipSuccess) {
            release();
            return fail(SPH_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", r.bytes, hipGetErrorString(e));
        }
        bytes += r.bytes;
    }
    b.list_cap = list_entries;
    if (hipHostMalloc((void **)&h_list, 8, 0) != hipSuccess) {
        h_list = nullptr;
        release();
        return fail(SPH_ERR_OUT_OF_MEMORY, "pinned host words");
    }
    h_list[0] = h_list[1] = 0u;
    hipError_t e = hipMemset(b.ghist, 0, kGhistCopies * kGhistWords * 4);
    if (e == hipSuccess) e = hipMemset(b.counters, 0, 8 * 4);
    if (e == hipSuccess) e = hipMemset(b.brick_flag, 0, (size_t)(brick_total ? brick_total : 1) * 4);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    brick_prepare();
    if (e != hipSuccess) {
        release();
        return hip_fail(e, "hipMemset");
    }
    return SPH_OK;
}

int Chain::set_debug(bool on) {
    if (on && !b.dbg_pred_vel) {
        hipError_t e = hipMalloc((void **)&b.dbg_pred_vel, (size_t)n * 16);
        if (e == hipSuccess) e = hipMalloc((void **)&b.dbg_vel_force, (size_t)n * 16);
        if (e != hipSuccess) return fail(SPH_ERR_OUT_OF_MEMORY, "debug buffers: %s", hipGetErrorString(e));
    } else if (!on && b.dbg_pred_vel) {
        hipFree(b.dbg_pred_vel);
        hipFree(b.dbg_vel_force);
        b.dbg_pred_vel = nullptr;
        b.dbg_vel_force = nullptr;
    }
    return SPH_OK;
}

int Chain::set_list_cap(uint32_t entries) {
    if (entries < 1u || entries > (uint32_t)kList)
        return fail(SPH_ERR_INVALID_ARG, "force lists of %u entries (1..%d)", entries, kList);
    if (entries == b.list_cap) return SPH_OK;
    uint32_t *fresh = nullptr;
    const size_t want = (size_t)n_cap * entries * 4;
    hipError_t e = hipMalloc((void **)&fresh, want);
    if (e != hipSuccess) return fail(SPH_ERR_OUT_OF_MEMORY, "force lists of %u entries: %s", entries, hipGetErrorString(e));
    if (b.nlist) (void)hipFree(b.nlist);
    bytes = bytes - (size_t)n_cap * b.list_cap * 4 + want;
    b.nlist = fresh;
    b.list_cap = entries;
    return SPH_OK;
}

uint32_t Chain::list_target(uint32_t record, bool shrink) {
    const uint32_t cap = b.list_cap;
    if (record > cap) {
        list_quiet = 0;
        if (cap >= (uint32_t)kList) return 0u;
        return std::min((uint32_t)kList, std::max(2u * cap, (record + record / 4u + 31u) / 32u * 32u));
    }
    if (record != 0u) {  // a list above half the capacity: keep it
        list_quiet = 0;
        return 0u;
    }
    if (++list_quiet < kListQuiet || cap <= kListStart || !shrink) return 0u;
    list_quiet = 0;
    return std::max(kListStart, (cap / 2u + 31u) / 32u * 32u);
}

int Chain::sync_lists(hipStream_t s, bool adapt) {
    adapt = adapt && h_list;
    // h_list[1]: this read; h_list[0]: the copy a step call may have queued (it cleared the
    // device record after copying, so both are needed)
    hipError_t e = adapt ? hipMemcpyAsync(h_list + 1, b.counters + 4, 4, hipMemcpyDeviceToHost, s) : hipSuccess;
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    const bool polled = list_poll_pending;
    list_poll_pending = false;  // any queued copy has landed
    if (!adapt) return 0;
    const uint32_t record = std::max(h_list[1], polled ? h_list[0] : 0u);
    if (h_list[1]) {
        // the stream is idle: clear the record for the next read
        e = hipMemsetAsync(b.counters + 4, 0, 4, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "clearing the list record");
    }
    const uint32_t want = list_target(record, true);
    return want && set_list_cap(want) == SPH_OK ? 1 : 0;
}

int Chain::poll_lists_async(hipStream_t s) {
    if (!h_list) return 0;
    if (!list_ev && hipEventCreateWithFlags(&list_ev, hipEventDisableTiming) != hipSuccess) {
        list_ev = nullptr;
        return 0;
    }
    if (list_poll_pending) {
        if (hipEventQuery(list_ev) != hipSuccess) return 0;  // not landed yet
        list_poll_pending = false;
        // growth only: a shrink waits for a synchronisation, where the stream is idle anyway
        // (here it would drain the queued frames and re-capture the graph mid-call, ~0.4 ms
        // on the 1M fill, for memory alone)
        const uint32_t want = list_target(h_list[0], false);
        if (want) {
            const hipError_t e = hipStreamSynchronize(s);  // the lists are free to reallocate
            if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
            return set_list_cap(want) == SPH_OK ? 1 : 0;
        }
    }
    // the copy, then the clear (stream-ordered: the next substeps record afresh); neither is
    // inside a captured graph
    if (hipMemcpyAsync(h_list, b.counters + 4, 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
        hipMemsetAsync(b.counters + 4, 0, 4, s) == hipSuccess && hipEventRecord(list_ev, s) == hipSuccess)
        list_poll_pending = true;
    return 0;
}

uint32_t list_entries_for_state(const sph_params &p, uint32_t n, const float *pos4) {
    // five times the neighbours of a uniform fill of the state's bounding box (each extent
    // at least one smoothing radius), in steps of 32, between kListStart and kList: a dense
    // start (a resting block, the reference's own scene) gets long lists at once instead of
    // walking until they grow.  Measured, the most neighbours a particle reaches in the
    // next frames (a random fill relaxing, a dam breaking, an over-dense block expanding)
    // stay within 4.9 times the mean at the start.
    float lo[3] = {pos4[0], pos4[1], pos4[2]}, hi[3] = {pos4[0], pos4[1], pos4[2]};
    for (uint32_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pos4[4 * (size_t)i + a]);
            hi[a] = std::max(hi[a], pos4[4 * (size_t)i + a]);
        }
    const double R = p.smoothing_radius;
    double vol = 1.0;
    for (int a = 0; a < 3; ++a) vol *= std::max((double)hi[a] - lo[a], R);
    const double neighbours = (double)n / vol * (4.0 / 3.0) * M_PI * R * R * R;
    const double want = std::ceil(5.0 * neighbours / 32.0) * 32.0;
    return (uint32_t)std::min<double>((double)kList, std::max<double>((double)kListStart, want));
}

double particles_per_cell_for_state(const sph_params &p, uint32_t n, const float *pos4) {
    // the mean over the cells of the state's bounding box, the box's grid at most
    float lo[3] = {pos4[0], pos4[1], pos4[2]}, hi[3] = {pos4[0], pos4[1], pos4[2]};
    for (uint32_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pos4[4 * (size_t)i + a]);
            hi[a] = std::max(hi[a], pos4[4 * (size_t)i + a]);
        }
    const double R = p.smoothing_radius;
    double cells = 1.0;
    for (int a = 0; a < 3; ++a) cells *= std::floor(((double)hi[a] - lo[a]) / R) + 1.0;
    return (double)n / cells;
}

uint32_t stray_blocks_for(uint32_t untiled, uint32_t n, int block) {
    // every particle of scenes without tiles, plus a sixteenth of the rest, in one sweep of
    // the stray workgroups (more strays take more sweeps)
    const uint64_t lanes = (uint64_t)untiled + ((uint64_t)n - untiled) / 16u;
    const uint64_t blocks = (lanes + (uint64_t)block - 1u) / (uint64_t)block;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(8u, blocks), 65535u);
}

void Chain::release() {
    void *ptrs[] = {b.pos, b.vel, b.keys, b.sort_k[0], b.sort_v[0], b.sort_k[1], b.sort_v[1],
                    b.pred_pos, b.pos_sorted, b.vel_sorted, b.dens, b.table, b.ghist,
                    b.lookback, b.tile_ctr, b.dbg_pred_vel, b.dbg_vel_force, b.hfull,
                    b.brick_flag, b.queue, b.stray, b.counters, b.bcount, b.blookback,
                    b.run_info, b.crun, b.hslot, b.grid, b.nlist, b.ncount, b.towner};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (h_list) (void)hipHostFree(h_list);
    h_list = nullptr;
    if (list_ev) (void)hipEventDestroy(list_ev);
    list_ev = nullptr;
    list_poll_pending = false;
    list_quiet = 0;
    b = ChainBuffers{};
    bytes = 0;
}

const uint32_t *Chain::sorted_keys() const { return passes ? b.sort_k[(passes - 1u) & 1u] : b.keys; }
const uint32_t *Chain::sorted_idx() const { return passes ? b.sort_v[(passes - 1u) & 1u] : nullptr; }

template <class Scenes>
hipError_t Chain::enqueue(hipStream_t s, const Scenes &sc, PassEvents *ev) {
    hipError_t e;
    if (ev) hipEventRecord(ev->ev[0], s);
    cell_order = !bricks;
    e = launch_hash(s, sc, b, n, table_total, passes, lookback_words, cell_order ? bucket_total : 0u, radix_bits);
    if (e != hipSuccess) return e;
    if (ev) hipEventRecord(ev->ev[1], s);
    const uint32_t bins = 1u << radix_bits;
    for (uint32_t q = 0; q < passes; ++q) {
        const uint32_t *kin = q == 0 ? b.keys : b.sort_k[(q - 1u) & 1u];
        const uint32_t *vin = q == 0 ? nullptr : b.sort_v[(q - 1u) & 1u];
        const uint32_t shift = radix_bits * q;
        const uint32_t bits = key_bits - shift < radix_bits ? key_bits - shift : radix_bits;
        e = radix_pass(s, kin, vin, b.sort_k[q & 1u], b.sort_v[q & 1u], n, shift, bits,
                       b.ghist + bins * q, ghist_words(), b.lookback + (size_t)q * tiles * bins, b.tile_ctr + q,
                       radix_bits);
        if (e != hipSuccess) return e;
    }
    if (ev) hipEventRecord(ev->ev[2], s);
    if ((e = launch_reorder(s, sc, b, sorted_keys(), sorted_idx(), n, bricks, ghist_words())) != hipSuccess) return e;
    if (bricks) {
        if ((e = launch_queue(s, b, brick_total)) != hipSuccess) return e;
    } else {
        if ((e = launch_bucket_scan(s, b, bucket_total)) != hipSuccess) return e;
        if ((e = launch_place(s, sc, b, sorted_keys(), sorted_idx(), n, grid_total, tiled())) != hipSuccess) return e;
    }
    if (ev) hipEventRecord(ev->ev[3], s);
    e = bricks ? launch_density_brick(s, sc, b, n, brick_total)
        : tiled() ? launch_density_tile(s, sc, b, n, tile_total, stray_blocks, tile_cfg)
                  : launch_density(s, sc, b, n, lane_cfg);
    if (e != hipSuccess) return e;
    if (ev) hipEventRecord(ev->ev[4], s);
    e = bricks ? launch_force_brick(s, sc, b, n, brick_total)
        : tiled() ? launch_force_tile(s, sc, b, n, tile_total, stray_blocks, tile_cfg)
                  : launch_force(s, sc, b, n, lane_cfg);
    if (e != hipSuccess) return e;
    if (ev) hipEventRecord(ev->ev[5], s);
    return hipSuccess;
}

template hipError_t Chain::enqueue<OneScene>(hipStream_t, const OneScene &, PassEvents *);
template hipError_t Chain::enqueue<ManyScenes>(hipStream_t, const ManyScenes &, PassEvents *);

int GraphCache::ensure_begin(hipStream_t s) {
    drop();
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "hipStreamBeginCapture");
}

int GraphCache::end(hipStream_t s, hipError_t launch_status) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (launch_status != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return hip_fail(launch_status, "kernel launch during capture");
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        exec = nullptr;
        return hip_fail(e, "hipGraphInstantiate");
    }
    return SPH_OK;
}

void GraphCache::drop() {
    if (exec) hipGraphExecDestroy(exec);
    exec = nullptr;
}

int PassEvents::create() {
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(SPH_ERR_DEVICE, "hipEventCreate");
    return SPH_OK;
}

void PassEvents::destroy() {
    for (auto &e : ev)
        if (e) hipEventDestroy(e);
}

void speed_ramp(uint32_t n, const float *vel4, float *color4) {
    static const float blue[3] = {0.3f, 0.3f, 1.0f}, cyan[3] = {0.0f, 1.0f, 1.0f},
                       yellow[3] = {1.0f, 1.0f, 0.0f}, red[3] = {1.0f, 0.3f, 0.3f};
    for (uint32_t i = 0; i < n; ++i) {
        const float *v = &vel4[4 * (size_t)i];
        const float speed = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const float t = std::fmin(std::fmax(speed / 60.0f, 0.0f), 1.0f);
        const float *a, *b;
        float w;
        if (t < 0.33f) { a = blue; b = cyan; w = t / 0.33f; }
        else if (t < 0.66f) { a = cyan; b = yellow; w = (t - 0.33f) / 0.33f; }
        else { a = yellow; b = red; w = (t - 0.66f) / 0.34f; }
        for (int q = 0; q < 3; ++q) color4[4 * (size_t)i + q] = a[q] * (1.0f - w) + b[q] * w;
        color4[4 * (size_t)i + 3] = 1.0f;
    }
}

}  // namespace sph

using namespace sph;

struct sph_world {
    sph_params params{};
    uint32_t n = 0;
    uint32_t T = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    Chain chain;
    OneScene scene{};
    GraphCache graph;
    bool debug = false;
    bool stepped = false;           // the table and intermediates match the state
    bool grid_current = true;       // the halo grid was filled with the current cell grid
    int variant = SPH_KERNEL_AUTO;  // requested neighbour-pass kernels
    bool tuned = false;             // AUTO resolved (by the tuning cache or a measurement)
    float *d_map = nullptr;         // density map staging for host copies
    // overlapped readback (sph_world_snapshot_async): state snapshot [2n] (positions, then
    // velocities), the stream that copies it to the host, and the two hand-over events
    float4 *d_snap = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_taken = nullptr, snap_copied = nullptr;
    bool snap_pending = false;
    size_t map_cap = 0;             // floats allocated at d_map
    // the world's own state buffers while the caller's are bound (sph_world_bind_state)
    float4 *own_pos = nullptr, *own_vel = nullptr;
    bool tolerance = false;         // SPH_PRECISION_TOLERANCE: the tuner may pick the
                                    // approximate-math configurations
    double ppc = 0.0;               // particles per occupied cell of the last uploaded state
                                    // (0: not known; the box's mean is used)
    uint32_t list_pinned = 0;       // force-list entries fixed by sph_world_set_list_capacity
                                    // (0: sized from the state and grown when lists overflow)
    // Online autotuning of an AUTO world: its first substeps run the candidate
    // configurations (two each, the second between events); all give bit-identical
    // results, so these are the world's real substeps and nothing is saved or restored.
    struct OnlineTune {
        bool active = false;
        int phase = 0;               // 0: kernel configurations, 1: radix widths
        std::vector<int> codes;
        std::vector<hipEvent_t> ev;  // [2 * code]: around the timed substep
        size_t next = 0;             // next code to run
        int sub = 0;                 // substeps run of the current code
        int best = -1;
        float best_ms = 0.0f;
    } online;
};

namespace {

// Autotuning: every kernel variant gives bit-identical results, so the choice only
// affects speed.  A variant is chosen per size bucket (particle count and table size,
// rounded to powers of two) by timing a few substeps of each candidate on the world's own
// state (saved and restored around the measurement) and cached for the process.
std::mutex g_tune_mu;
std::map<uint64_t, int> g_tune_cache;

uint64_t tune_bucket(uint32_t n, uint32_t T, uint64_t bricks, bool tolerance) {
    return ((uint64_t)key_bits_for(n) << 48) | ((uint64_t)key_bits_for(T) << 32) |
           ((uint64_t)tolerance << 31) | (uint64_t)key_bits_for(bricks ? bricks : 1);
}

// Launch configurations the world may use at its precision.
int lane_configs_for(const sph_world *w) { return w->tolerance ? kNumLaneConfigs : kNumExactLaneConfigs; }

// A tuning code names one configuration: 0..kNumLaneConfigs-1 = the per-lane kernels with
// that launch configuration, kCodeBrick = the brick kernels; bits 8 and up carry the radix
// digit width (0: 8 bits).
constexpr int kCodeBrick = 100;
constexpr int kCodeTile = 200;  // kCodeTile + c: the tiled kernels at tile configuration c
constexpr int kCodeMask = 0xFF;

int tile_configs_for(const sph_world *w) { return w->tolerance ? kNumTileConfigs : kNumExactTileConfigs; }

// Tiles of the world's scene at tile configuration cfg (-1: the per-lane kernels), sized
// for the particles per occupied cell of the last uploaded state.
bool world_tiles_possible(const sph_world *w);

void world_set_tiles(sph_world *w, int cfg) {
    SceneConst &c = w->scene.c;
    if (cfg < 0 || !world_tiles_possible(w)) {
        c.use_tiles = 0u;
        w->chain.tile_cfg = -1;
        w->chain.tile_total = 0;
        return;
    }
    const double ppc = w->ppc > 0.0 ? w->ppc : (double)w->n / ((double)c.gx * c.gy * c.gz);
    tile_dims_for(c, ppc, cfg);
    c.use_tiles = 1u;
    c.tile_base = 0u;
    w->chain.tile_cfg = cfg;
    w->chain.tile_total = scene_tiles(c);
    w->chain.stray_blocks = stray_blocks_for(0, w->n, 256);
}

// Tiles cover the box's cell grid: at most 64 cells per particle (+ 65536), so that empty
// tiles stay cheap; cells without a halo grid are looked up through the hash table.
bool world_tiles_possible(const sph_world *w) {
    const SceneConst &c = w->scene.c;
    return (double)c.gx * c.gy * c.gz <= 64.0 * w->n + 65536.0;
}

int world_apply_code(sph_world *w, int code) {
    const int kernels = code & kCodeMask;
    w->chain.bricks = kernels == kCodeBrick && w->chain.brick_total > 0;
    if (kernels >= 0 && kernels < kNumLaneConfigs) w->chain.lane_cfg = kernels;
    world_set_tiles(w, kernels >= kCodeTile && kernels < kCodeTile + kNumTileConfigs ? kernels - kCodeTile : -1);
    const uint32_t width = (uint32_t)code >> 8;
    if (!w->chain.set_radix_bits(width ? width : 8u)) w->chain.set_radix_bits(8u);
    w->graph.drop();
    return SPH_OK;
}

// Synchronises the world's stream; an adaptive world (lists not pinned) grows its force
// lists to hold the longest that did not fit.  Returns 1 when they grew, 0, or an error.
int world_sync_lists(sph_world *w) {
    const int rc = w->chain.sync_lists(w->stream, !w->list_pinned);
    if (rc == 1) w->graph.drop();
    return rc;
}

// Times `substeps` eager substeps of each configuration; restores the state afterwards.
int world_autotune(sph_world *w, uint32_t substeps, int *chosen, double *ms_out) {
    const size_t bytes = (size_t)w->n * 16;
    float4 *save_p = nullptr, *save_v = nullptr;
    hipError_t e = hipMalloc((void **)&save_p, bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&save_v, bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(save_p, w->chain.b.pos, bytes, hipMemcpyDeviceToDevice, w->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(save_v, w->chain.b.vel, bytes, hipMemcpyDeviceToDevice, w->stream);
    if (e != hipSuccess) {
        if (save_p) (void)hipFree(save_p);
        if (save_v) (void)hipFree(save_v);
        return hip_fail(e, "autotune state copy");
    }
    std::vector<int> cands;
    for (int c = 0; c < lane_configs_for(w); ++c) cands.push_back(c);
    if (w->chain.brick_total) cands.push_back(kCodeBrick);
    if (world_tiles_possible(w))
        for (int c = 0; c < tile_configs_for(w); ++c)
            if (tile_config_tuned(c, w->tolerance)) cands.push_back(kCodeTile + c);
    int best = kDefaultLaneConfig;
    double best_ms = 1e30;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    // kernels first (at 8-bit digits), then the radix widths with the chosen kernels
    const size_t n_kernel_cands = cands.size();
    for (size_t ci = 0;; ++ci) {
        if (ci == n_kernel_cands && e == hipSuccess) {
            // widths that need fewer passes than 8-bit digits, with the fastest kernels
            const int kbest = best & kCodeMask;
            for (uint32_t width : {11u, 12u})
                if (w->chain.lookback_words_for(width) <= w->chain.lookback_cap &&
                    (w->chain.key_bits + width - 1u) / width < (w->chain.key_bits + 7u) / 8u)
                    cands.push_back(kbest | (int)(width << 8));
        }
        if (ci >= cands.size()) break;
        const int code = cands[ci];
        world_apply_code(w, code);
        // one untimed substep, then the timed ones, each from the saved state
        e = hipMemcpyAsync(w->chain.b.pos, save_p, bytes, hipMemcpyDeviceToDevice, w->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(w->chain.b.vel, save_v, bytes, hipMemcpyDeviceToDevice, w->stream);
        if (e == hipSuccess) e = w->chain.enqueue(w->stream, w->scene, nullptr);
        if (e == hipSuccess) e = hipEventRecord(t0, w->stream);
        for (uint32_t s = 0; s < substeps && e == hipSuccess; ++s) e = w->chain.enqueue(w->stream, w->scene, nullptr);
        if (e == hipSuccess) e = hipEventRecord(t1, w->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        if (e != hipSuccess) break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (ms < best_ms) {
            best_ms = ms;
            best = code;
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(w->chain.b.pos, save_p, bytes, hipMemcpyDeviceToDevice, w->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(w->chain.b.vel, save_v, bytes, hipMemcpyDeviceToDevice, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(save_p);
    (void)hipFree(save_v);
    if (e != hipSuccess) return hip_fail(e, "autotune");
    w->stepped = false;  // the table is the last trial's, not the restored state's
    world_apply_code(w, best);
    *chosen = best;
    if (ms_out) *ms_out = best_ms / substeps;
    return SPH_OK;
}

void online_release(sph_world *w) {
    for (hipEvent_t e : w->online.ev)
        if (e) (void)hipEventDestroy(e);
    w->online.ev.clear();
}

// Ends online tuning without choosing (an explicit choice or measurement replaces it).
void online_cancel(sph_world *w) {
    if (!w->online.active) return;
    (void)hipStreamSynchronize(w->stream);
    online_release(w);
    w->online.active = false;
}

int online_begin_phase(sph_world *w, std::vector<int> codes) {
    online_release(w);
    w->online.codes = std::move(codes);
    w->online.ev.assign(2 * w->online.codes.size(), nullptr);
    for (hipEvent_t &e : w->online.ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(SPH_ERR_DEVICE, "hipEventCreate");
    w->online.next = 0;
    w->online.sub = 0;
    return SPH_OK;
}

int online_start(sph_world *w) {
    std::vector<int> codes;
    for (int c = 0; c < lane_configs_for(w); ++c) codes.push_back(c);
    if (w->chain.brick_total) codes.push_back(kCodeBrick);
    if (world_tiles_possible(w))
        for (int c = 0; c < tile_configs_for(w); ++c)
            if (tile_config_tuned(c, w->tolerance)) codes.push_back(kCodeTile + c);
    w->online = sph_world::OnlineTune{};
    w->online.active = true;
    return online_begin_phase(w, std::move(codes));
}

// All codes of the phase ran: wait for them (work already queued), keep the fastest; after
// the kernels, try the radix widths that need fewer passes with the fastest kernels.
int online_finish_phase(sph_world *w) {
    auto &o = w->online;
    const uint32_t cap_before = w->chain.b.list_cap;
    const int resized = world_sync_lists(w);
    if (resized < 0) return resized;
    if (w->chain.b.list_cap > cap_before && o.phase == 0) {
        // the list configurations ran with lists too short for this state: time them again
        std::vector<int> codes = o.codes;
        return online_begin_phase(w, std::move(codes));
    }
    for (size_t i = 0; i < o.codes.size(); ++i) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, o.ev[2 * i], o.ev[2 * i + 1]) != hipSuccess) continue;
        if (o.best < 0 || ms < o.best_ms) {
            o.best = o.codes[i];
            o.best_ms = ms;
        }
    }
    if (o.best < 0) o.best = kDefaultLaneConfig;
    if (o.phase == 0) {
        std::vector<int> widths;
        const int kbest = o.best & kCodeMask;
        for (uint32_t width : {11u, 12u})
            if (w->chain.lookback_words_for(width) <= w->chain.lookback_cap &&
                (w->chain.key_bits + width - 1u) / width < (w->chain.key_bits + 7u) / 8u)
                widths.push_back(kbest | (int)(width << 8));
        if (!widths.empty()) {
            o.phase = 1;
            return online_begin_phase(w, std::move(widths));
        }
    }
    online_release(w);
    o.active = false;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[tune_bucket(w->n, w->T, w->chain.brick_total, w->tolerance)] = o.best;
    }
    world_apply_code(w, o.best);
    w->tuned = true;
    return SPH_OK;
}

// One substep of online tuning.
hipError_t online_substep(sph_world *w) {
    auto &o = w->online;
    if (o.sub == 0) {
        world_apply_code(w, o.codes[o.next]);
        o.sub = 1;
        return w->chain.enqueue(w->stream, w->scene, nullptr);
    }
    hipError_t e = hipEventRecord(o.ev[2 * o.next], w->stream);
    if (e == hipSuccess) e = w->chain.enqueue(w->stream, w->scene, nullptr);
    if (e == hipSuccess) e = hipEventRecord(o.ev[2 * o.next + 1], w->stream);
    o.sub = 0;
    ++o.next;
    return e;
}

// Resolves AUTO before the first step of a configuration: the tuning cache, a pinned
// configuration, or a measurement (online over the coming substeps when `online` is set,
// else now, on a saved copy of the state).
int world_resolve_variant(sph_world *w, bool online = false) {
    if (w->tuned || w->variant != SPH_KERNEL_AUTO) return SPH_OK;
    if (w->online.active) {
        if (online) return SPH_OK;
        online_cancel(w);
    }
    const uint64_t key = tune_bucket(w->n, w->T, w->chain.brick_total, w->tolerance);
    int code = -1;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) code = it->second;
    }
    if (code < 0) {
        const char *off = std::getenv("SPH_AUTOTUNE");
        const char *pin = std::getenv("SPH_LANE_CONFIG");
        const char *tpin = std::getenv("SPH_TILE_CONFIG");
        const int pinned = pin ? std::atoi(pin) : -1;
        const int tpinned = tpin ? std::atoi(tpin) : -1;
        if (tpin && *tpin >= '0' && *tpin <= '9' && tpinned < tile_configs_for(w) && world_tiles_possible(w)) {
            code = kCodeTile + tpinned;  // a fixed tiled configuration (profiling)
        } else if (pin && *pin >= '0' && *pin <= '9' && pinned >= 0 && pinned < lane_configs_for(w)) {
            code = pinned;  // a fixed per-lane launch configuration (profiling)
        } else if (off && std::strcmp(off, "0") == 0) {
            code = kDefaultLaneConfig;
        } else if (online) {
            return online_start(w);
        } else {
            int rc = world_autotune(w, 3, &code, nullptr);
            if (rc) return rc;
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[key] = code;
    }
    world_apply_code(w, code);
    w->tuned = true;
    return SPH_OK;
}

}  // namespace

extern "C" {

void sph_params_derive(sph_params *p, double frame_dt, uint32_t substeps, double smoothing_radius) {
    if (!p || substeps == 0) return;
    p->substeps = substeps;
    p->dt = (float)(frame_dt / (double)substeps);
    p->volume = (float)(2.0 * M_PI * std::pow(smoothing_radius, 5) / 15.0);
}

uint64_t sph_world_footprint(uint32_t n, uint32_t table_cells) {
    const uint64_t T = table_cells ? table_cells : n;
    const uint64_t tiles = ((uint64_t)n + radix_tile_size() - 1) / radix_tile_size();
    // state, keys, 2 x radix ping-pong, 3 sorted float4 arrays, densities, brick and
    // cell-order slot arrays, force lists at their largest capacity (kList entries); per
    // table cell its (start, end), run placement and cell-ordered run; buckets (at most
    // 2n + 1024), the halo grid (at most 16n + 4096 cells) and look-back words
    return (uint64_t)n * (16 * 2 + 4 + 16 + 16 * 3 + 8 + 8 + 4 + 4 * kList + 1) + T * 24 +
           (2 * (uint64_t)n + 1024) * 4 + (16 * (uint64_t)n + 4096) * 8 +
           std::max<uint64_t>(4 * tiles * 256 * 4, std::max<uint64_t>(16 * (uint64_t)n, 8ull << 20)) +
           kGhistCopies * kGhistWords * 4 + 4096;
}

int sph_world_create(const sph_params *params, uint32_t n, sph_world **out) {
    if (!out) return fail(SPH_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int rc = validate_params(params);
    if (rc) return rc;
    if (n == 0) return fail(SPH_ERR_INVALID_ARG, "particle count must be > 0");
    if (n >= (1u << 30)) return fail(SPH_ERR_CAPACITY, "particle count must be < 2^30");
    const uint32_t T = params->table_cells ? params->table_cells : n;
    if (T >= (1u << 31)) return fail(SPH_ERR_CAPACITY, "table_cells must be < 2^31");
    sph_world *w = new (std::nothrow) sph_world();
    if (!w) return fail(SPH_ERR_OUT_OF_MEMORY, "host allocation");
    w->params = *params;
    w->params.table_cells = T;
    w->n = n;
    w->T = T;
    hipError_t e = hipGetDevice(&w->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete w;
        return hip_fail(e, "stream creation");
    }
    w->scene.c = make_scene_const(w->params, T, 0, 0, n, 0, 0, 0);
    rc = w->chain.allocate(n, T, (uint32_t)scene_bricks(w->scene.c), scene_buckets(w->scene.c),
                           w->scene.c.grid_cells, kListStart);
    if (rc) {
        hipStreamDestroy(w->stream);
        delete w;
        return rc;
    }
    world_apply_code(w, kDefaultLaneConfig);
    *out = w;
    return SPH_OK;
}

int sph_world_set_params(sph_world *w, const sph_params *params) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    int rc = validate_params(params);
    if (rc) return rc;
    const uint32_t T = params->table_cells ? params->table_cells : w->n;
    if (T != w->T) return fail(SPH_ERR_INVALID_ARG, "table_cells is fixed at creation");
    SceneConst nc = make_scene_const(*params, T, 0, 0, w->n, 0, 0, 0);
    if (scene_bricks(nc) != w->chain.brick_total || scene_buckets(nc) != w->chain.bucket_total ||
        nc.grid_cells != w->chain.grid_total)
        return fail(SPH_ERR_INVALID_ARG, "the box and smoothing radius fix the cell grid at creation");
    // equal totals may still lay the grid out differently (e.g. two box extents swapped):
    // until the next substep refills it, the halo grid does not match the new dimensions
    if (nc.gx != w->scene.c.gx || nc.gy != w->scene.c.gy || nc.gz != w->scene.c.gz) w->grid_current = false;
    w->params = *params;
    w->params.table_cells = T;
    w->scene.c = nc;
    world_set_tiles(w, w->chain.tile_cfg);
    w->graph.drop();
    return SPH_OK;
}

int sph_world_upload_state(sph_world *w, const float *pos4, const float *vel4) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    int rc = validate_state(w->params, w->n, pos4, vel4);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(w->chain.b.pos, pos4, (size_t)w->n * 16, hipMemcpyHostToDevice, w->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(w->chain.b.vel, vel4, (size_t)w->n * 16, hipMemcpyHostToDevice, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "upload");
    w->stepped = false;
    if (!w->list_pinned) {
        const uint32_t want = list_entries_for_state(w->params, w->n, pos4);
        if (want > w->chain.b.list_cap && w->chain.set_list_cap(want) == SPH_OK) w->graph.drop();
    }
    w->ppc = particles_per_cell_for_state(w->params, w->n, pos4);
    if (w->chain.tile_cfg >= 0) {
        const SceneConst before = w->scene.c;
        world_set_tiles(w, w->chain.tile_cfg);
        if (before.tw != w->scene.c.tw || before.th != w->scene.c.th || before.td != w->scene.c.td) w->graph.drop();
    }
    return SPH_OK;
}

int sph_world_step(sph_world *w, uint32_t frames) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    const uint64_t substeps = (uint64_t)frames * w->params.substeps;
    if (substeps == 0) return SPH_OK;
    int rc0 = world_resolve_variant(w, true);
    if (rc0) return rc0;
    for (uint64_t s = 0; s < substeps; ++s) {
        if (w->online.active && w->online.next == w->online.codes.size()) {
            int rc = online_finish_phase(w);
            if (rc) return rc;
        }
        if (w->online.active) {
            hipError_t e = online_substep(w);
            if (e != hipSuccess) return hip_fail(e, "autotuning substep");
            continue;
        }
        if (!w->graph.exec) {
            int rc = w->graph.ensure_begin(w->stream);
            if (rc) return rc;
            hipError_t le = w->chain.enqueue(w->stream, w->scene, nullptr);
            rc = w->graph.end(w->stream, le);
            if (rc) return rc;
        }
        hipError_t e = hipGraphLaunch(w->graph.exec, w->stream);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
        // once a frame, without waiting: the lists resize as the last landed record asks
        if (!w->list_pinned && (s + 1) % w->params.substeps == 0) {
            const int rc = w->chain.poll_lists_async(w->stream);
            if (rc < 0) return rc;
            if (rc == 1) w->graph.drop();
        }
    }
    w->stepped = true;
    w->grid_current = true;
    return SPH_OK;
}

int sph_world_sync(sph_world *w) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    const int rc = world_sync_lists(w);
    return rc < 0 ? rc : SPH_OK;
}

int sph_world_download_state(sph_world *w, float *pos4, float *vel4, float *color4) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    std::vector<float> tmp;
    float *vdst = vel4;
    if (color4 && !vel4) {
        tmp.resize((size_t)w->n * 4);
        vdst = tmp.data();
    }
    hipError_t e = hipSuccess;
    if (pos4) e = hipMemcpyAsync(pos4, w->chain.b.pos, (size_t)w->n * 16, hipMemcpyDeviceToHost, w->stream);
    if (e == hipSuccess && vdst)
        e = hipMemcpyAsync(vdst, w->chain.b.vel, (size_t)w->n * 16, hipMemcpyDeviceToHost, w->stream);
    if (e != hipSuccess) return hip_fail(e, "download");
    const int rc = world_sync_lists(w);
    if (rc < 0) return rc;
    if (color4) speed_ramp(w->n, vdst, color4);
    return SPH_OK;
}

int sph_world_snapshot_wait(sph_world *w) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (!w->snap_pending) return SPH_OK;
    const hipError_t e = hipEventSynchronize(w->snap_copied);
    w->snap_pending = false;
    return e == hipSuccess ? SPH_OK : hip_fail(e, "snapshot wait");
}

int sph_world_snapshot_async(sph_world *w, float *pos4, float *vel4) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    int rc = sph_world_snapshot_wait(w);  // the snapshot buffer is free again
    if (rc) return rc;
    hipError_t e = hipSuccess;
    if (!w->copy_stream) {
        e = hipMalloc(&w->d_snap, (size_t)w->n * 32);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&w->copy_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&w->snap_taken, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&w->snap_copied, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "snapshot set-up");
    }
    const size_t bytes = (size_t)w->n * 16;
    float4 *snap_vel = w->d_snap + w->n;
    if (pos4) e = hipMemcpyAsync(w->d_snap, w->chain.b.pos, bytes, hipMemcpyDeviceToDevice, w->stream);
    if (e == hipSuccess && vel4)
        e = hipMemcpyAsync(snap_vel, w->chain.b.vel, bytes, hipMemcpyDeviceToDevice, w->stream);
    if (e == hipSuccess) e = hipEventRecord(w->snap_taken, w->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(w->copy_stream, w->snap_taken, 0);
    if (e == hipSuccess && pos4) e = hipMemcpyAsync(pos4, w->d_snap, bytes, hipMemcpyDeviceToHost, w->copy_stream);
    if (e == hipSuccess && vel4) e = hipMemcpyAsync(vel4, snap_vel, bytes, hipMemcpyDeviceToHost, w->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(w->snap_copied, w->copy_stream);
    if (e != hipSuccess) return hip_fail(e, "snapshot");
    w->snap_pending = true;
    return SPH_OK;
}

int sph_world_set_kernel(sph_world *w, int variant) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (variant < SPH_KERNEL_AUTO || variant > SPH_KERNEL_TILE)
        return fail(SPH_ERR_INVALID_ARG, "unknown kernel variant %d", variant);
    if (variant == SPH_KERNEL_BRICK && w->chain.brick_total == 0)
        return fail(SPH_ERR_INVALID_ARG, "brick kernels need a grid of at most 2^28 bricks");
    if (variant == SPH_KERNEL_TILE && !world_tiles_possible(w))
        return fail(SPH_ERR_INVALID_ARG, "tiled kernels need a box of at most 64n + 65536 cells");
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    online_cancel(w);
    w->variant = variant;
    w->tuned = variant != SPH_KERNEL_AUTO;
    const int kernels = variant == SPH_KERNEL_BRICK ? kCodeBrick
                      : variant == SPH_KERNEL_TILE   ? kCodeTile + (w->chain.tile_cfg >= 0 ? w->chain.tile_cfg : 0)
                                                     : w->chain.lane_cfg;
    world_apply_code(w, kernels | (int)(w->chain.radix_bits == 8u ? 0u : w->chain.radix_bits << 8));
    return SPH_OK;
}

int sph_world_set_tile_config(sph_world *w, int config) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (config < 0 || config >= kNumTileConfigs) return fail(SPH_ERR_INVALID_ARG, "tile configuration %d", config);
    if (kTileConfigs[config].tolerance && !w->tolerance)
        return fail(SPH_ERR_INVALID_ARG, "tile configuration %d needs SPH_PRECISION_TOLERANCE", config);
    if (!world_tiles_possible(w))
        return fail(SPH_ERR_INVALID_ARG, "tiled kernels need a box of at most 64n + 65536 cells");
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    online_cancel(w);
    w->variant = SPH_KERNEL_TILE;
    w->tuned = true;
    world_apply_code(w, (kCodeTile + config) | (int)(w->chain.radix_bits == 8u ? 0u : w->chain.radix_bits << 8));
    return SPH_OK;
}

int sph_tile_configs(void) { return kNumTileConfigs; }

int sph_world_tile_stats(sph_world *w, uint32_t out[4]) {
    if (!w || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    uint32_t c[8];
    hipError_t e = hipMemcpyAsync(c, w->chain.b.counters, sizeof c, hipMemcpyDeviceToHost, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "tile statistics");
    const bool tiled = w->chain.tiled() && w->stepped;
    out[0] = tiled ? w->chain.tile_total : 0u;
    out[1] = tiled ? c[1] + c[7] : 0u;
    out[2] = tiled ? c[5] : 0u;
    out[3] = tiled ? c[6] : 0u;
    return SPH_OK;
}

int sph_world_autotune(sph_world *w, uint32_t substeps, int *chosen, double *ms_per_substep) {
    if (!w || !chosen) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (substeps == 0) return fail(SPH_ERR_INVALID_ARG, "substeps must be >= 1");
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    online_cancel(w);
    int code = 0;
    int rc = 0;
    for (;;) {
        rc = world_autotune(w, substeps, &code, ms_per_substep);
        if (rc) return rc;
        // the list configurations' trials overflowed lists that have now grown: again
        const uint32_t cap_before = w->chain.b.list_cap;
        const int resized = world_sync_lists(w);
        if (resized < 0) return resized;
        if (w->chain.b.list_cap <= cap_before) break;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[tune_bucket(w->n, w->T, w->chain.brick_total, w->tolerance)] = code;
    w->variant = SPH_KERNEL_AUTO;
    w->tuned = true;
    *chosen = (code & kCodeMask) == kCodeBrick ? SPH_KERNEL_BRICK
            : (code & kCodeMask) >= kCodeTile  ? SPH_KERNEL_TILE
                                               : SPH_KERNEL_DIRECT;
    return SPH_OK;
}

int sph_world_set_lane_config(sph_world *w, int config) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (config < 0 || config >= kNumLaneConfigs)
        return fail(SPH_ERR_INVALID_ARG, "lane configuration %d outside [0, %d)", config, kNumLaneConfigs);
    if (kLaneConfigs[config].tolerance && !w->tolerance)
        return fail(SPH_ERR_STATE, "lane configuration %d is not bit-exact: select SPH_PRECISION_TOLERANCE first", config);
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    online_cancel(w);
    w->variant = SPH_KERNEL_DIRECT;
    w->tuned = true;
    world_apply_code(w, config | (int)(w->chain.radix_bits == 8u ? 0u : w->chain.radix_bits << 8));
    return SPH_OK;
}

int sph_world_set_radix_bits(sph_world *w, int bits) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (bits != 8 && bits != 11 && bits != 12) return fail(SPH_ERR_INVALID_ARG, "radix digits of %d bits: use 8, 11 or 12", bits);
    if (w->chain.lookback_words_for((uint32_t)bits) > w->chain.lookback_cap)
        return fail(SPH_ERR_CAPACITY, "%d-bit radix digits need more look-back memory than this world holds", bits);
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    w->chain.set_radix_bits((uint32_t)bits);
    w->graph.drop();
    if (w->variant == SPH_KERNEL_AUTO) {  // keep the tuned kernels, pin this width
        int rc = world_resolve_variant(w);
        if (rc) return rc;
        w->chain.set_radix_bits((uint32_t)bits);
        w->variant = w->chain.bricks ? SPH_KERNEL_BRICK : SPH_KERNEL_DIRECT;
    }
    return SPH_OK;
}

int sph_lane_configs(void) { return kNumLaneConfigs; }

int sph_world_set_list_capacity(sph_world *w, uint32_t entries) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (entries > (uint32_t)kList) return fail(SPH_ERR_INVALID_ARG, "force lists hold at most %d entries", kList);
    w->list_pinned = entries;
    if (entries == 0) return SPH_OK;
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    const int rc = w->chain.set_list_cap(entries);
    if (rc) return rc;
    w->graph.drop();
    return SPH_OK;
}

int sph_world_list_capacity(const sph_world *w) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    return (int)w->chain.b.list_cap;
}

uint64_t sph_world_device_bytes(const sph_world *w) {
    if (!w) return 0;
    return w->chain.bytes + (w->d_snap ? (uint64_t)w->n * 32 : 0) + (uint64_t)w->map_cap * 4;
}

int sph_world_set_precision(sph_world *w, int mode) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (mode != SPH_PRECISION_EXACT && mode != SPH_PRECISION_TOLERANCE)
        return fail(SPH_ERR_INVALID_ARG, "unknown precision mode %d", mode);
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    online_cancel(w);
    w->tolerance = mode == SPH_PRECISION_TOLERANCE;
    if (!w->tolerance && kLaneConfigs[w->chain.lane_cfg].tolerance) {
        world_apply_code(w, kDefaultLaneConfig);
        w->variant = SPH_KERNEL_AUTO;
    }
    if (w->variant == SPH_KERNEL_AUTO) w->tuned = false;  // tune again for this precision
    return SPH_OK;
}

int sph_world_kernel(const sph_world *w) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    return w->chain.bricks ? SPH_KERNEL_BRICK : w->chain.tiled() ? SPH_KERNEL_TILE : SPH_KERNEL_DIRECT;
}

int sph_world_kernel_config(const sph_world *w, char *buf, size_t bytes) {
    if (!w || !buf || bytes == 0) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    char radix[64];
    snprintf(radix, sizeof radix, "%u-bit radix digits (%u passes)", w->chain.radix_bits, w->chain.passes);
    if (w->chain.bricks)
        snprintf(buf, bytes, "brick %dx%dx%d cells, persistent queue, %s", kBrick, kBrick, kBrick, radix);
    else if (w->chain.tiled())
        snprintf(buf, bytes, "tiles of %ux%ux%u cells staged in LDS (%d-lane workgroups, %d entries), %s%s",
                 w->scene.c.tw, w->scene.c.th, w->scene.c.td, kTileConfigs[w->chain.tile_cfg].block,
                 kTileConfigs[w->chain.tile_cfg].entries, radix,
                 kTileConfigs[w->chain.tile_cfg].tolerance ? ", approximate square root and reciprocal" : "");
    else
        snprintf(buf, bytes, "per-lane, %d candidates per scan step, %d-lane workgroups, %s, %s%s",
                 kLaneConfigs[w->chain.lane_cfg].unroll, kLaneConfigs[w->chain.lane_cfg].block,
                 kLaneConfigs[w->chain.lane_cfg].lists ? "force neighbour lists" : "force cell walk", radix,
                 kLaneConfigs[w->chain.lane_cfg].tolerance ? ", approximate square root and reciprocal" : "");
    return SPH_OK;
}

int sph_world_set_debug(sph_world *w, int enable) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    int rc = w->chain.set_debug(enable != 0);
    if (rc) return rc;
    w->debug = enable != 0;
    w->graph.drop();
    return SPH_OK;
}

int sph_world_debug_dump(sph_world *w, int stage, void *dst, size_t bytes) {
    if (!w || !dst) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (!w->stepped) return fail(SPH_ERR_STATE, "no substep has run since the last upload");
    return dump_stage(w->chain, w->stream, stage, 0, w->n, 0, w->T, dst, bytes);
}

int sph_world_pass_times(sph_world *w, uint32_t substeps, sph_pass_times *out) {
    if (!w || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (substeps == 0) return fail(SPH_ERR_INVALID_ARG, "substeps must be >= 1");
    int rc = world_resolve_variant(w);
    if (rc) return rc;
    PassEvents ev;
    rc = ev.create();
    if (rc) { ev.destroy(); return rc; }
    double acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t s = 0; s < substeps; ++s) {
        hipError_t e = w->chain.enqueue(w->stream, w->scene, &ev);
        if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
        if (e != hipSuccess) { ev.destroy(); return hip_fail(e, "timed substep"); }
        for (int q = 0; q < 5; ++q) {
            float ms = 0.0f;
            hipEventElapsedTime(&ms, ev.ev[q], ev.ev[q + 1]);
            acc[q] += ms;
        }
    }
    ev.destroy();
    out->hash_ms = acc[0] / substeps;
    out->sort_ms = acc[1] / substeps;
    out->reorder_ms = acc[2] / substeps;
    out->density_ms = acc[3] / substeps;
    out->force_ms = acc[4] / substeps;
    out->total_ms = out->hash_ms + out->sort_ms + out->reorder_ms + out->density_ms + out->force_ms;
    w->stepped = true;
    w->grid_current = true;
    return SPH_OK;
}

int sph_world_density_map_dims(const sph_world *w, float voxel_size, uint32_t dims[3]) {
    if (!w || !dims) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    MapGeometry g{};
    int rc = map_geometry(w->params, voxel_size, g);
    if (rc) return rc;
    dims[0] = g.dx;
    dims[1] = g.dy;
    dims[2] = g.dz;
    return SPH_OK;
}

int sph_density_map_dims(const sph_params *params, float voxel_size, uint32_t dims[3]) {
    if (!params || !dims) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    int rc = validate_params(params);
    if (rc) return rc;
    MapGeometry g{};
    rc = map_geometry(*params, voxel_size, g);
    if (rc) return rc;
    dims[0] = g.dx;
    dims[1] = g.dy;
    dims[2] = g.dz;
    return SPH_OK;
}

int sph_world_density_map_device(sph_world *w, float voxel_size, float *dev_out, size_t count) {
    if (!w || !dev_out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (!w->stepped) return fail(SPH_ERR_STATE, "no substep has run since the last upload");
    MapGeometry g{};
    int rc = map_geometry(w->params, voxel_size, g);
    if (rc) return rc;
    if (count < g.voxels) return fail(SPH_ERR_INVALID_ARG, "density map needs %llu floats, got %zu",
                                      (unsigned long long)g.voxels, count);
    hipError_t e = launch_density_map(w->stream, w->scene.c, w->chain, w->grid_current, g, dev_out);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "density map launch");
}

int sph_world_density_map(sph_world *w, float voxel_size, float *out, size_t count) {
    if (!w || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    MapGeometry g{};
    int rc = map_geometry(w->params, voxel_size, g);
    if (rc) return rc;
    if (count < g.voxels) return fail(SPH_ERR_INVALID_ARG, "density map needs %llu floats, got %zu",
                                      (unsigned long long)g.voxels, count);
    if (!w->stepped) return fail(SPH_ERR_STATE, "no substep has run since the last upload");
    if (w->map_cap < g.voxels) {
        hipError_t e = hipStreamSynchronize(w->stream);
        if (e == hipSuccess && w->d_map) e = hipFree(w->d_map);
        w->d_map = nullptr;
        w->map_cap = 0;
        if (e == hipSuccess) e = hipMalloc((void **)&w->d_map, g.voxels * sizeof(float));
        if (e != hipSuccess) return fail(SPH_ERR_OUT_OF_MEMORY, "density map buffer: %s", hipGetErrorString(e));
        w->map_cap = g.voxels;
    }
    hipError_t e = launch_density_map(w->stream, w->scene.c, w->chain, w->grid_current, g, w->d_map);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, w->d_map, g.voxels * sizeof(float), hipMemcpyDeviceToHost, w->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(w->stream);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "density map");
}

int sph_world_bind_state(sph_world *w, float *pos4_dev, float *vel4_dev) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if ((pos4_dev == nullptr) != (vel4_dev == nullptr))
        return fail(SPH_ERR_INVALID_ARG, "bind both state buffers or neither");
    if (pos4_dev && pos4_dev == vel4_dev) return fail(SPH_ERR_INVALID_ARG, "positions and velocities alias");
    const uintptr_t mis = (uintptr_t)pos4_dev | (uintptr_t)vel4_dev;
    if (mis & 15u) return fail(SPH_ERR_INVALID_ARG, "state buffers must be 16-byte aligned");
    int rc = sph_world_snapshot_wait(w);
    if (rc) return rc;
    hipError_t e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if (pos4_dev) {
        hipPointerAttribute_t a{}, b{};
        e = hipPointerGetAttributes(&a, pos4_dev);
        if (e == hipSuccess) e = hipPointerGetAttributes(&b, vel4_dev);
        if (e != hipSuccess || a.type == hipMemoryTypeHost || b.type == hipMemoryTypeHost ||
            a.type == hipMemoryTypeUnregistered || b.type == hipMemoryTypeUnregistered) {
            (void)hipGetLastError();
            return fail(SPH_ERR_INVALID_ARG, "state buffers must be device memory");
        }
        if (!w->own_pos) {
            w->own_pos = w->chain.b.pos;
            w->own_vel = w->chain.b.vel;
        }
        w->chain.b.pos = reinterpret_cast<float4 *>(pos4_dev);
        w->chain.b.vel = reinterpret_cast<float4 *>(vel4_dev);
    } else if (w->own_pos) {
        w->chain.b.pos = w->own_pos;
        w->chain.b.vel = w->own_vel;
        w->own_pos = w->own_vel = nullptr;
    }
    w->graph.drop();  // the captured launches name the previous buffers
    w->stepped = false;
    return SPH_OK;
}

int sph_world_device_state(const sph_world *w, float **pos4_dev, float **vel4_dev) {
    if (!w) return fail(SPH_ERR_INVALID_ARG, "world is NULL");
    if (pos4_dev) *pos4_dev = reinterpret_cast<float *>(w->chain.b.pos);
    if (vel4_dev) *vel4_dev = reinterpret_cast<float *>(w->chain.b.vel);
    return SPH_OK;
}

void sph_world_destroy(sph_world *w) {
    if (!w) return;
    if (w->stream) hipStreamSynchronize(w->stream);
    online_release(w);
    if (w->own_pos) {  // never free the caller's buffers
        w->chain.b.pos = w->own_pos;
        w->chain.b.vel = w->own_vel;
    }
    if (w->copy_stream) {
        hipStreamSynchronize(w->copy_stream);
        hipStreamDestroy(w->copy_stream);
    }
    if (w->snap_taken) hipEventDestroy(w->snap_taken);
    if (w->snap_copied) hipEventDestroy(w->snap_copied);
    if (w->d_snap) hipFree(w->d_snap);
    if (w->d_map) hipFree(w->d_map);
    w->graph.drop();
    w->chain.release();
    if (w->stream) hipStreamDestroy(w->stream);
    delete w;
}

const char *sph_last_error(void) { return sph::last_error(); }

}  // extern "C"
