// Batched step: many independent scenes of uneven size in one chain of launches.
//
// Scene s owns particle slots [p_begin_s, p_begin_s + n_s) and table cells
// [t_begin_s, t_begin_s + T_s).  Its cell keys are t_begin_s + (hash & (T_s - 1)), i.e. the
// scene id sits above the local key bits of one segmented key space, so a single stable
// radix sort orders every scene's particles inside its own slot range exactly as a sort of
// that scene alone would.  The hash, reorder and table kernels look a slot's scene up in a
// 2-byte scene-id array; the density and force kernels drain one work queue that holds the
// bricks of every dense scene followed by lane chunks of the sparse ones.  One hipGraph per
// substep covers all scenes.
//
// Each particle's arithmetic sees only its own scene's constants and runs, in the same
// order as in a single-scene world, so every scene's result is bit-identical alone and in
// any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/sph_world.h"
#include "sph_map.hpp"
#include "sph_host.hpp"
#include "sph_kernels.hpp"

using namespace sph;

struct sph_batch {
    uint32_t count = 0;
    std::vector<sph_params> params;
    std::vector<uint32_t> n, T, p_begin, t_begin;
    std::vector<SceneConst> scenes;
    std::vector<uint8_t> brick_choice;  // per scene: 0 auto, 1 direct, 2 brick
    uint32_t n_total = 0, t_total = 0, b_total = 0, substeps = 1;
    int variant = SPH_KERNEL_AUTO;
    int tuned_variant = -1;  // AUTO's measured choice; -1 = measure before the next step
    hipStream_t stream = nullptr;
    Chain chain;
    SceneConst *d_scenes = nullptr;
    uint16_t *d_scene_of = nullptr;
    GraphCache graph;
    bool stepped = false;
    float *d_map = nullptr;  // density map staging for host copies
    size_t map_cap = 0;
    uint32_t list_pinned = 0;  // force-list entries fixed by sph_batch_set_list_capacity (0: adaptive)
    std::vector<double> ppc;   // per scene: particles per cell of its last uploaded state (0: unknown)
    int tile_cfg = 0;          // tile configuration TILE uses
    ManyScenes policy() const { return ManyScenes{d_scenes, d_scene_of, count}; }
};

namespace sph {

// scene_of[j] = the scene whose slot range [p_begin, p_end) holds slot j.
__global__ __launch_bounds__(256) void k_fill_scene_of(const SceneConst *__restrict__ table,
                                                       uint32_t count, uint16_t *__restrict__ scene_of,
                                                       uint32_t n) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= n) return;
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (table[mid].p_begin <= j) lo = mid; else hi = mid;
    }
    scene_of[j] = (uint16_t)lo;
}

hipError_t launch_fill_scene_of(hipStream_t stream, const SceneConst *table, uint32_t count,
                                uint16_t *scene_of, uint32_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_scene_of, dim3((n + 255u) / 256u), dim3(256), 0, stream, table, count,
                       scene_of, n);
    return hipGetLastError();
}

}  // namespace sph

namespace {

constexpr int kBatchCodeTile = 200;  // autotuner code of tile configuration c: kBatchCodeTile + c

// A scene runs through the brick kernels when its particles are dense enough for a brick's
// halo staging to be reused: at least 0.25 particles per cell of the box its particles
// span (16 per 4x4x4 brick).
constexpr double kBrickDensity = 0.25;

uint32_t choose_bricks(const SceneConst &c, uint32_t n, const float *pos4, int variant) {
    if (variant == SPH_KERNEL_DIRECT) return 0u;
    if (variant == SPH_KERNEL_BRICK) return 1u;
    if (!pos4) return 1u;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], pos4[4 * (size_t)i + a]);
            hi[a] = std::max(hi[a], pos4[4 * (size_t)i + a]);
        }
    double cells = 1.0;
    for (int a = 0; a < 3; ++a) cells *= std::floor((hi[a] - lo[a]) / c.R) + 1.0;
    return (double)n / cells >= kBrickDensity ? 1u : 0u;
}

int upload_scene_table(sph_batch *b) {
    hipError_t e = hipMemcpyAsync(b->d_scenes, b->scenes.data(), b->scenes.size() * sizeof(SceneConst),
                                  hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "scene table upload");
}

}  // namespace

extern "C" {

int sph_batch_create(uint32_t count, const sph_params *params, const uint32_t *counts,
                     sph_batch **out) {
    if (!out || !params || !counts) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (count == 0 || count > 65535u) return fail(SPH_ERR_INVALID_ARG, "scene count must be in [1, 65535]");
    uint64_t n_total = 0, t_total = 0, b_total = 0;
    for (uint32_t s = 0; s < count; ++s) {
        int rc = validate_params(&params[s]);
        if (rc) return fail(rc, "scene %u: %s", s, last_error());
        if (counts[s] == 0) return fail(SPH_ERR_INVALID_ARG, "scene %u has no particles", s);
        if (params[s].substeps != params[0].substeps)
            return fail(SPH_ERR_INVALID_ARG, "scene %u: every scene of a batch needs the same substep count", s);
        n_total += counts[s];
        t_total += params[s].table_cells ? params[s].table_cells : counts[s];
    }
    if (n_total >= (1ull << 30)) return fail(SPH_ERR_CAPACITY, "batch exceeds 2^30 particles");
    if (t_total >= (1ull << 31)) return fail(SPH_ERR_CAPACITY, "batch exceeds 2^31 table cells");
    sph_batch *b = new (std::nothrow) sph_batch();
    if (!b) return fail(SPH_ERR_OUT_OF_MEMORY, "host allocation");
    b->count = count;
    b->params.assign(params, params + count);
    b->substeps = params[0].substeps;
    uint32_t pb = 0, tb = 0, bb = 0, ub = 0, gb = 0;
    for (uint32_t s = 0; s < count; ++s) {
        const uint32_t T = params[s].table_cells ? params[s].table_cells : counts[s];
        b->params[s].table_cells = T;
        b->n.push_back(counts[s]);
        b->T.push_back(T);
        b->p_begin.push_back(pb);
        b->t_begin.push_back(tb);
        SceneConst c = make_scene_const(b->params[s], T, tb, pb, pb + counts[s], bb, ub, gb);
        ub += scene_buckets(c);  // <= 2 n_s + 1024 each: below 2^31 + 2^26 in total
        if ((uint64_t)gb + c.grid_cells >= (1ull << 31)) c.grid_cells = 0u;  // hash lookups only
        gb += c.grid_cells;
        const uint64_t nb = scene_bricks(c);
        if (nb == 0 || b_total + nb >= (1ull << 31)) c.use_bricks = 0u;  // stays on lanes
        else { bb += (uint32_t)nb; b_total += nb; }
        b->scenes.push_back(c);
        pb += counts[s];
        tb += T;
    }
    b->n_total = (uint32_t)n_total;
    b->t_total = (uint32_t)t_total;
    b->b_total = (uint32_t)b_total;
    b->brick_choice.assign(count, 0);
    b->ppc.assign(count, 0.0);
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return hip_fail(e, "stream creation");
    }
    int rc = b->chain.allocate(b->n_total, b->t_total, b->b_total ? b->b_total : 1u, ub, gb, kListStart);
    if (rc) {
        hipStreamDestroy(b->stream);
        delete b;
        return rc;
    }
    b->chain.bricks = false;  // AUTO starts on the per-lane kernels until tuned
    std::vector<uint16_t> scene_of(b->n_total);
    for (uint32_t s = 0; s < count; ++s)
        std::fill(scene_of.begin() + b->p_begin[s], scene_of.begin() + b->p_begin[s] + b->n[s], (uint16_t)s);
    e = hipMalloc((void **)&b->d_scenes, count * sizeof(SceneConst));
    if (e == hipSuccess) e = hipMalloc((void **)&b->d_scene_of, (size_t)b->n_total * 2);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_scene_of, scene_of.data(), (size_t)b->n_total * 2, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        sph_batch_destroy(b);
        return fail(SPH_ERR_OUT_OF_MEMORY, "scene tables: %s", hipGetErrorString(e));
    }
    if ((rc = upload_scene_table(b))) {
        sph_batch_destroy(b);
        return rc;
    }
    *out = b;
    return SPH_OK;
}

// Tiles for every scene whose box has at most 64 cells per particle (+ 65536), sized for its
// last uploaded state, at tile configuration cfg (-1: none); the other scenes' particles are
// the tiled launches' strays.
static void batch_set_tiles(sph_batch *b, int cfg) {
    uint32_t tile_base = 0, untiled = 0;
    for (uint32_t s = 0; s < b->count; ++s) {
        SceneConst &c = b->scenes[s];
        const double cells = (double)c.gx * c.gy * c.gz;
        c.use_tiles = 0u;
        c.tile_base = tile_base;
        if (cfg >= 0 && cells <= 64.0 * b->n[s] + 65536.0) {
            tile_dims_for(c, b->ppc[s] > 0.0 ? b->ppc[s] : (double)b->n[s] / cells, cfg);
            const uint64_t t = (uint64_t)c.ntx * c.nty * c.ntz;
            if ((uint64_t)tile_base + t < (1ull << 31)) {
                c.use_tiles = 1u;
                tile_base += (uint32_t)t;
            }
        }
        if (!c.use_tiles) untiled += b->n[s];
    }
    b->chain.tile_cfg = cfg >= 0 && tile_base > 0 ? cfg : -1;
    b->chain.tile_total = tile_base;
    b->chain.stray_blocks = stray_blocks_for(untiled, b->n_total, 256);
}

// DIRECT: every scene on the per-lane kernels.  BRICK: the queue kernels, with bricks for
// the scenes whose brick grid exists and whose density (AUTO's per-scene choice, or all of
// them when `all` is set) makes staging pay; the others ride along as lane chunks.  TILE:
// the tiled kernels at tile configuration `tile_cfg`.
static int batch_apply(sph_batch *b, int v, bool all, int tile_cfg = 0) {
    for (uint32_t s = 0; s < b->count; ++s) {
        uint32_t ub = 0u;
        if (v == SPH_KERNEL_BRICK && b->b_total && scene_bricks(b->scenes[s]))
            ub = all ? 1u : b->brick_choice[s];
        b->scenes[s].use_bricks = ub;
    }
    b->chain.bricks = v == SPH_KERNEL_BRICK && b->b_total > 0;
    batch_set_tiles(b, v == SPH_KERNEL_TILE ? tile_cfg : -1);
    b->graph.drop();
    return upload_scene_table(b);
}

// Synchronises the batch's stream; adaptive force lists grow to hold the longest one that
// did not fit (one capacity for every scene).  Returns 1 when they grew, 0, or an error.
static int batch_sync_lists(sph_batch *b) {
    const int rc = b->chain.sync_lists(b->stream, !b->list_pinned);
    if (rc == 1) b->graph.drop();
    return rc;
}

// Times a few substeps of each configuration on a saved copy of the state and keeps the
// fastest (results are identical either way).
static int batch_autotune_once(sph_batch *b, uint32_t substeps) {
    const size_t bytes = (size_t)b->n_total * 16;
    float4 *sp = nullptr, *sv = nullptr;
    hipError_t e = hipMalloc((void **)&sp, bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&sv, bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(sp, b->chain.b.pos, bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(sv, b->chain.b.vel, bytes, hipMemcpyDeviceToDevice, b->stream);
    // Candidates: the per-lane kernels in each launch configuration (code = config index),
    // then the brick mix (code = -1).
    int best = kDefaultLaneConfig;
    float best_ms = 3.0e38f;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    // Then, with the fastest kernels, the radix widths that need fewer passes than 8-bit
    // digits (code = kernels | width << 8).
    std::vector<int> cands;
    for (int c = 0; c < kNumExactLaneConfigs; ++c) cands.push_back(c);
    if (b->b_total) cands.push_back(-1);
    for (int c = 0; c < kNumExactTileConfigs; ++c)
        if (tile_config_tuned(c, false)) cands.push_back(kBatchCodeTile + c);
    b->chain.set_radix_bits(8u);
    const size_t n_kernel_cands = cands.size();
    for (size_t ci = 0;; ++ci) {
        if (ci == n_kernel_cands && e == hipSuccess && best >= 0) {
            const uint32_t kb = b->chain.key_bits;
            for (uint32_t width : {11u, 12u})
                if (b->chain.lookback_words_for(width) <= b->chain.lookback_cap &&
                    (kb + width - 1u) / width < (kb + 7u) / 8u)
                    cands.push_back(best | (int)(width << 8));
        }
        if (ci >= cands.size() || e != hipSuccess) break;
        const int code = cands[ci];
        const int v = code < 0 ? -1 : (code & 0xFF);
        b->chain.set_radix_bits(code < 0 || (code >> 8) == 0 ? 8u : (uint32_t)(code >> 8));
        if (v >= 0 && v < kBatchCodeTile) b->chain.lane_cfg = v;
        int rc = batch_apply(b, v < 0 ? SPH_KERNEL_BRICK : v >= kBatchCodeTile ? SPH_KERNEL_TILE : SPH_KERNEL_DIRECT,
                             false, v - kBatchCodeTile);
        if (rc) { e = hipErrorUnknown; break; }
        e = hipMemcpyAsync(b->chain.b.pos, sp, bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(b->chain.b.vel, sv, bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e == hipSuccess) e = b->chain.enqueue(b->stream, b->policy(), nullptr);
        if (e == hipSuccess) e = hipEventRecord(t0, b->stream);
        for (uint32_t k = 0; k < substeps && e == hipSuccess; ++k) e = b->chain.enqueue(b->stream, b->policy(), nullptr);
        if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (e == hipSuccess && ms < best_ms) { best_ms = ms; best = code; }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(b->chain.b.pos, sp, bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->chain.b.vel, sv, bytes, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (sp) (void)hipFree(sp);
    if (sv) (void)hipFree(sv);
    if (e != hipSuccess) return hip_fail(e, "batch autotune");
    b->stepped = false;  // the table is the last trial's, not the restored state's
    const int bv = best < 0 ? -1 : (best & 0xFF);
    b->tuned_variant = bv < 0 ? SPH_KERNEL_BRICK : bv >= kBatchCodeTile ? SPH_KERNEL_TILE : SPH_KERNEL_DIRECT;
    if (bv >= 0 && bv < kBatchCodeTile) b->chain.lane_cfg = bv;
    b->tile_cfg = bv >= kBatchCodeTile ? bv - kBatchCodeTile : b->tile_cfg;
    b->chain.set_radix_bits(best < 0 || (best >> 8) == 0 ? 8u : (uint32_t)(best >> 8));
    return batch_apply(b, b->tuned_variant, false, b->tile_cfg);
}

// The measurement again while the trials grew the force lists (their list configurations
// were timed with lists too short for the state).
static int batch_autotune(sph_batch *b, uint32_t substeps) {
    for (;;) {
        int rc = batch_autotune_once(b, substeps);
        if (rc) return rc;
        const uint32_t cap_before = b->chain.b.list_cap;
        const int resized = batch_sync_lists(b);
        if (resized < 0) return resized;
        if (b->chain.b.list_cap <= cap_before) return SPH_OK;
    }
}

int sph_batch_set_kernel(sph_batch *b, int variant) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    if (variant < SPH_KERNEL_AUTO || variant > SPH_KERNEL_TILE)
        return fail(SPH_ERR_INVALID_ARG, "unknown kernel variant %d", variant);
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    b->variant = variant;
    b->tuned_variant = variant == SPH_KERNEL_AUTO ? -1 : variant;
    return batch_apply(b, variant == SPH_KERNEL_AUTO ? SPH_KERNEL_DIRECT : variant, variant == SPH_KERNEL_BRICK,
                       b->tile_cfg);
}

int sph_batch_set_tile_config(sph_batch *b, int config) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    if (config < 0 || config >= kNumExactTileConfigs)
        return fail(SPH_ERR_INVALID_ARG, "tile configuration %d (a batch takes the exact ones)", config);
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    b->variant = SPH_KERNEL_TILE;
    b->tuned_variant = SPH_KERNEL_TILE;
    b->tile_cfg = config;
    return batch_apply(b, SPH_KERNEL_TILE, false, config);
}

int sph_batch_upload_scene(sph_batch *b, uint32_t scene, const float *pos4, const float *vel4) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    if (scene >= b->count) return fail(SPH_ERR_INVALID_ARG, "scene %u out of range", scene);
    int rc = validate_state(b->params[scene], b->n[scene], pos4, vel4);
    if (rc) return rc;
    const size_t off = b->p_begin[scene], bytes = (size_t)b->n[scene] * 16;
    hipError_t e = hipMemcpyAsync(b->chain.b.pos + off, pos4, bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->chain.b.vel + off, vel4, bytes, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "upload");
    b->brick_choice[scene] = (uint8_t)choose_bricks(b->scenes[scene], b->n[scene], pos4, SPH_KERNEL_AUTO);
    b->ppc[scene] = particles_per_cell_for_state(b->params[scene], b->n[scene], pos4);
    if (b->chain.tile_cfg >= 0) {
        batch_set_tiles(b, b->chain.tile_cfg);  // the scene's tiles follow its new density
        b->graph.drop();
        if ((rc = upload_scene_table(b))) return rc;
    }
    if (!b->list_pinned) {
        const uint32_t want = list_entries_for_state(b->params[scene], b->n[scene], pos4);
        if (want > b->chain.b.list_cap && b->chain.set_list_cap(want) == SPH_OK) b->graph.drop();
    }
    if (b->variant == SPH_KERNEL_AUTO) b->tuned_variant = -1;  // re-tune for the new state
    b->stepped = false;
    return SPH_OK;
}

int sph_batch_step(sph_batch *b, uint32_t frames) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    const uint64_t substeps = (uint64_t)frames * b->substeps;
    if (substeps == 0) return SPH_OK;
    if (b->tuned_variant < 0) {
        int rc = batch_autotune(b, 2);
        if (rc) return rc;
    }
    for (uint64_t s = 0; s < substeps; ++s) {
        if (!b->graph.exec) {
            int rc = b->graph.ensure_begin(b->stream);
            if (rc) return rc;
            hipError_t le = b->chain.enqueue(b->stream, b->policy(), nullptr);
            rc = b->graph.end(b->stream, le);
            if (rc) return rc;
        }
        hipError_t e = hipGraphLaunch(b->graph.exec, b->stream);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
        // once a frame, without waiting: the lists resize as the last landed record asks
        if (!b->list_pinned && (s + 1) % b->substeps == 0) {
            const int rc = b->chain.poll_lists_async(b->stream);
            if (rc < 0) return rc;
            if (rc == 1) b->graph.drop();
        }
    }
    b->stepped = true;
    return SPH_OK;
}

int sph_batch_sync(sph_batch *b) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    const int rc = batch_sync_lists(b);
    return rc < 0 ? rc : SPH_OK;
}

int sph_batch_set_list_capacity(sph_batch *b, uint32_t entries) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    if (entries > (uint32_t)kList) return fail(SPH_ERR_INVALID_ARG, "force lists hold at most %d entries", kList);
    b->list_pinned = entries;
    if (entries == 0) return SPH_OK;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    const int rc = b->chain.set_list_cap(entries);
    if (rc) return rc;
    b->graph.drop();
    return SPH_OK;
}

int sph_batch_list_capacity(const sph_batch *b) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    return (int)b->chain.b.list_cap;
}

int sph_batch_download_scene(sph_batch *b, uint32_t scene, float *pos4, float *vel4, float *color4) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    if (scene >= b->count) return fail(SPH_ERR_INVALID_ARG, "scene %u out of range", scene);
    const size_t off = b->p_begin[scene], n = b->n[scene];
    std::vector<float> tmp;
    float *vdst = vel4;
    if (color4 && !vel4) {
        tmp.resize(n * 4);
        vdst = tmp.data();
    }
    hipError_t e = hipSuccess;
    if (pos4) e = hipMemcpyAsync(pos4, b->chain.b.pos + off, n * 16, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess && vdst)
        e = hipMemcpyAsync(vdst, b->chain.b.vel + off, n * 16, hipMemcpyDeviceToHost, b->stream);
    if (e != hipSuccess) return hip_fail(e, "download");
    const int rc = batch_sync_lists(b);
    if (rc < 0) return rc;
    if (color4) speed_ramp((uint32_t)n, vdst, color4);
    return SPH_OK;
}

int sph_batch_set_debug(sph_batch *b, int enable) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    int rc = b->chain.set_debug(enable != 0);
    if (rc) return rc;
    b->graph.drop();
    return SPH_OK;
}

int sph_batch_debug_dump(sph_batch *b, uint32_t scene, int stage, void *dst, size_t bytes) {
    if (!b || !dst) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (scene >= b->count) return fail(SPH_ERR_INVALID_ARG, "scene %u out of range", scene);
    if (!b->stepped) return fail(SPH_ERR_STATE, "no substep has run since the last upload");
    return dump_stage(b->chain, b->stream, stage, b->p_begin[scene], b->p_begin[scene] + b->n[scene],
                      b->t_begin[scene], b->t_begin[scene] + b->T[scene], dst, bytes);
}

int sph_batch_kernel(const sph_batch *b) {
    if (!b) return fail(SPH_ERR_INVALID_ARG, "batch is NULL");
    return b->tuned_variant < 0 ? SPH_KERNEL_AUTO : b->tuned_variant;
}

int sph_batch_scene_uses_bricks(sph_batch *b, uint32_t scene) {
    if (!b || scene >= b->count) return fail(SPH_ERR_INVALID_ARG, "bad batch or scene");
    return b->chain.bricks && b->scenes[scene].use_bricks ? 1 : 0;
}

int sph_batch_pass_times(sph_batch *b, uint32_t substeps, sph_pass_times *out) {
    if (!b || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (substeps == 0) return fail(SPH_ERR_INVALID_ARG, "substeps must be >= 1");
    int rc = SPH_OK;
    if (b->tuned_variant < 0 && (rc = batch_autotune(b, 2))) return rc;
    PassEvents ev;
    rc = ev.create();
    if (rc) { ev.destroy(); return rc; }
    double acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t s = 0; s < substeps; ++s) {
        hipError_t e = b->chain.enqueue(b->stream, b->policy(), &ev);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
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
    b->stepped = true;
    return SPH_OK;
}

uint32_t sph_batch_particles(const sph_batch *b) { return b ? b->n_total : 0u; }

int sph_batch_density_map(sph_batch *b, uint32_t scene, float voxel_size, float *out, size_t count) {
    if (!b || !out) return fail(SPH_ERR_INVALID_ARG, "NULL argument");
    if (scene >= b->count) return fail(SPH_ERR_INVALID_ARG, "scene %u out of range", scene);
    MapGeometry g{};
    int rc = map_geometry(b->params[scene], voxel_size, g);
    if (rc) return rc;
    if (count < g.voxels) return fail(SPH_ERR_INVALID_ARG, "density map needs %llu floats, got %zu",
                                      (unsigned long long)g.voxels, count);
    if (!b->stepped) return fail(SPH_ERR_STATE, "no substep has run since the last upload");
    if (b->map_cap < g.voxels) {
        hipError_t e = hipStreamSynchronize(b->stream);
        if (e == hipSuccess && b->d_map) e = hipFree(b->d_map);
        b->d_map = nullptr;
        b->map_cap = 0;
        if (e == hipSuccess) e = hipMalloc((void **)&b->d_map, g.voxels * sizeof(float));
        if (e != hipSuccess) return fail(SPH_ERR_OUT_OF_MEMORY, "density map buffer: %s", hipGetErrorString(e));
        b->map_cap = g.voxels;
    }
    hipError_t e = launch_density_map(b->stream, b->scenes[scene], b->chain, true, g, b->d_map);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, b->d_map, g.voxels * sizeof(float), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "density map");
}

void sph_batch_destroy(sph_batch *b) {
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->d_map) hipFree(b->d_map);
    b->graph.drop();
    b->chain.release();
    if (b->d_scenes) hipFree(b->d_scenes);
    if (b->d_scene_of) hipFree(b->d_scene_of);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
