// SPH substep kernels for gfx950: hash, reorder + cell table, density, force + integrate.
//
// One lane per particle slot; 256-lane workgroups (4 wave64) for the hash and reorder, an
// autotuned workgroup size (64, 128 or 256 lanes) for the neighbour passes.  Neighbour sums visit the
// 27 cells in the reference's (dz, dy, dx) order and each cell's run in ascending slot
// order, so every particle's floating-point sum has the reference's association order and
// does not depend on launch geometry or on which other scenes share the launch.
#include "sph_device.hpp"
#include "sph_kernels.hpp"
#include "sph_pairs.hpp"


namespace sph {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kHashBlock = 1024;      // lanes of a hashing / clearing workgroup
constexpr uint32_t kHashKeys = 4 * 1024;  // keys per hashing workgroup

// ---------------------------------------------------------------------------------------
// Hash: updateCellEntries.comp:69-80 (key of the CURRENT position, value = slot index,
// implicit here), plus the digit histograms of the radix passes and the clears that
// clearStartIndices.comp:47-52 performs, moved to a point where the table is dead.
template <class Scenes, int kW>
__global__ __launch_bounds__(kHashBlock) void k_hash(
    Scenes sc, const float4 *__restrict__ pos, uint32_t *__restrict__ keys, uint32_t n,
    uint32_t passes, uint32_t *__restrict__ ghist, uint32_t hash_blocks,
    int2 *__restrict__ table, int2 *__restrict__ crun, uint32_t table_total,
    uint32_t *__restrict__ lookback, uint32_t lookback_words, uint32_t *__restrict__ tile_ctr,
    uint32_t *__restrict__ counters,
    uint32_t *__restrict__ bcount, uint32_t bucket_total, uint32_t *__restrict__ blookback,
    uint32_t blookback_words) {
    constexpr uint32_t kBins = 1u << kW;         // radix digits of kW bits
    constexpr uint32_t kMaxP = (32u + kW - 1u) / kW;
    __shared__ uint32_t s_hist[kMaxP][kBins];
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x < hash_blocks) {
        for (uint32_t i = tid; i < passes * kBins; i += kHashBlock) (&s_hist[0][0])[i] = 0u;
        __syncthreads();
        const uint32_t per_block = ((n + hash_blocks - 1u) / hash_blocks + kHashBlock - 1u) / kHashBlock * kHashBlock;
        const uint32_t begin = blockIdx.x * per_block;
        const uint32_t end = min(n, begin + per_block);
        // The state arrives in the previous substep's key order, so a wave's digits come in
        // runs of one value: each run adds its length with one LDS atomic from its first
        // lane, where one atomic per key would queue the workgroup's 16 waves on one address.
        const uint32_t lane = tid & 63u;
        for (uint32_t i0 = begin; i0 < end; i0 += kHashBlock) {
            const uint32_t i = i0 + tid;
            const bool valid = i < end;
            uint32_t key = 0u;
            if (valid) {
                const SceneConst c = sc.scene(i);
                const float4 p = pos[i];
                const uint32_t cx = (uint32_t)cell_of(p.x, c.minx, c.R);
                const uint32_t cy = (uint32_t)cell_of(p.y, c.miny, c.R);
                const uint32_t cz = (uint32_t)cell_of(p.z, c.minz, c.R);
                key = c.table_base + (cell_hash(cx, cy, cz) & c.mask);
                keys[i] = key;
            }
            for (uint32_t q = 0; q < passes; ++q) {
                // lanes past the end carry a value no digit has
                const uint32_t d = valid ? (key >> (kW * q)) & (kBins - 1u) : 0xFFFFFFFFu;
                const uint32_t up = __shfl_up(d, 1u, 64);
                const bool head = lane == 0u || up != d;
                const uint64_t heads = __ballot(head);
                if (head && valid) {
                    const uint64_t later = lane == 63u ? 0ull : heads >> (lane + 1u);
                    const uint32_t len = later ? (uint32_t)__builtin_ctzll(later) + 1u : 64u - lane;
                    atomicAdd(&s_hist[q][d], len);
                }
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < passes * kBins; i += kHashBlock) {
            const uint32_t v = (&s_hist[0][0])[i];
            if (v) atomicAdd(&ghist[(blockIdx.x % kGhistCopies) * passes * kBins + i], v);
        }
    } else {
        const uint32_t b = blockIdx.x - hash_blocks;
        const uint32_t nb = gridDim.x - hash_blocks;
        if (b == 0 && tid < kTileCounters) tile_ctr[tid] = 0u;
        // [0] queue, [1] strays, [5] streamed tiles, [6] handed-over tiles, [7] their
        // particles (not [4]: the host clears the list record when it reads it)
        if (b == 0 && counters && (tid < 2 || (tid >= 5 && tid < 8))) counters[tid] = 0u;
        for (uint32_t i = b * kHashBlock + tid; i < table_total; i += nb * kHashBlock) {
            table[i] = make_int2(-1, -1);
            if (crun) crun[i] = make_int2(-1, -1);
        }
        for (uint32_t i = b * kHashBlock + tid; i < lookback_words; i += nb * kHashBlock) lookback[i] = 0u;
        for (uint32_t i = b * kHashBlock + tid; i < bucket_total; i += nb * kHashBlock) bcount[i] = 0u;
        for (uint32_t i = b * kHashBlock + tid; i < blookback_words; i += nb * kHashBlock) blookback[i] = 0u;
    }
}

// ---------------------------------------------------------------------------------------
// Reorder: updateStartIndices.comp:40-64 marks the run boundaries.  On the brick path it
// also does what scatterPass.comp:50-65 does, gathering the state into key order, and
// evaluates predictPositions.comp (a pure per-particle function) on the gathered state.
template <class Scenes>
__global__ __launch_bounds__(kBlock) void k_reorder(
    Scenes sc, const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ sidx,
    uint32_t n, const float4 *__restrict__ pos, const float4 *__restrict__ vel,
    float4 *__restrict__ pos_sorted, float4 *__restrict__ vel_sorted,
    float4 *__restrict__ pred_pos, float4 *__restrict__ dbg_pred_vel, int2 *__restrict__ table,
    uint32_t *__restrict__ ghist, uint32_t *__restrict__ hfull, uint32_t *__restrict__ brick_flag,
    uint32_t *__restrict__ stray, uint32_t *__restrict__ counters, uint32_t *__restrict__ bcount,
    uint2 *__restrict__ run_info, uint32_t ghist_words) {
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x == 0)
        for (uint32_t i = tid; i < kGhistCopies * ghist_words; i += kBlock) ghist[i] = 0u;
    const uint32_t j = blockIdx.x * kBlock + tid;
    bool to_stray = false;
    if (j < n) {
        const uint32_t key = skeys[j];
        const uint32_t prev = j ? skeys[j - 1u] : 0u;
        const bool head = j == 0 || key != prev;
        int *tbl = reinterpret_cast<int *>(table);
        if (head) {
            tbl[2u * key] = (int)j;
            if (j) tbl[2u * prev + 1u] = (int)j;
        }
        if (j == n - 1u) tbl[2u * key + 1u] = (int)n;
        // Cell-ordered layout: the run's bucket (its first particle's cell) and its offset in
        // the bucket.  A run ends at the next run head of the wave (one ballot); only the
        // wave's last run, which may continue past the wave, scans the keys beyond it.
        const uint64_t heads = bcount ? __ballot(head) : 0ull;
        if (bcount && head) {
            const uint32_t lane = tid & 63u;
            const uint64_t above = heads & ~((2ull << lane) - 1ull);
            uint32_t e;
            if (above) {
                e = j - lane + (uint32_t)(__ffsll((long long)above) - 1);
            } else {
                e = min(n, j - lane + 64u);
                while (e < n && skeys[e] == key) ++e;
            }
            const SceneConst c = sc.scene(j);
            const float4 P = pos[sidx ? sidx[j] : j];
            const uint32_t bucket = bucket_of(c, cell_of(P.x, c.minx, c.R), cell_of(P.y, c.miny, c.R),
                                              cell_of(P.z, c.minz, c.R));
            run_info[key] = make_uint2(bucket, atomicAdd(&bcount[bucket], e - j));
        }
        if (brick_flag) {
            const uint32_t idx = sidx ? sidx[j] : j;
            const float4 P = pos[idx];
            const float4 V = vel[idx];
            const SceneConst c = sc.scene(j);
            float3 p = make_float3(P.x, P.y, P.z), v = make_float3(V.x, V.y, V.z);
            kick_drift<false>(c, p, v);
            pos_sorted[j] = P;
            vel_sorted[j] = V;
            pred_pos[j] = make_float4(p.x, p.y, p.z, 1.0f);
            if (dbg_pred_vel) dbg_pred_vel[j] = make_float4(v.x, v.y, v.z, 1.0f);
            // Work for the brick kernels: the full hash of the particle's current cell
            // tells its own cell's run apart from colliding cells; its brick is marked
            // non-empty.  Particles of one-lane-per-particle scenes and particles outside
            // the brick grid go to the stray list instead.
            const int32_t cx = cell_of(P.x, c.minx, c.R), cy = cell_of(P.y, c.miny, c.R),
                          cz = cell_of(P.z, c.minz, c.R);
            hfull[j] = cell_hash((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
            if (c.use_bricks && cx >= 0 && cy >= 0 && cz >= 0 && cx < c.gx && cy < c.gy && cz < c.gz) {
                const uint32_t b = c.brick_base + (uint32_t)(cx / kBrick) +
                                   c.nbx * ((uint32_t)(cy / kBrick) + c.nby * (uint32_t)(cz / kBrick));
                brick_flag[b] = 1u;
            } else {
                to_stray = true;
            }
        }
    }
    if (brick_flag) {
        // One atomic per wave keeps a wave's slots consecutive in the stray list.
        const uint64_t m = __ballot(to_stray);
        if (m) {
            const int lane = tid & 63;
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&counters[1], (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (to_stray) stray[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = j;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Density: calculateDensities.comp:96-155, one lane per particle, runs read from global.
// Slots, pred_pos and `table` (the run table) are in cell order.
// kU: candidates loaded per scan iteration; kB: workgroup size (tuned per size bucket).
// Run lookup of the cell-ordered layout for scene c.
__device__ __forceinline__ GridRuns grid_runs(const SceneConst &c, const ChainBuffers &b) {
    return GridRuns{b.grid, c.grid_base, (uint32_t)c.gx + 2u, (uint32_t)c.gy + 2u,
                    c.grid_cells ? (uint32_t)c.gz + 2u : 0u, c.table_base, c.mask, b.crun};
}

// Workgroups are handed to the 8 XCDs round-robin.  xcd_chunked deals them in chunks of
// kChunk consecutive blocks instead: XCD x runs chunks x, x + 8, x + 16, ... (each chunk a
// contiguous slot range, whose neighbouring cells share that XCD's L2), the blocks of a
// last partial round of chunks in launch order.  One contiguous eighth of the slots per
// XCD puts the dense floor and wall layers on a few XCDs; launch order loses the L2 reuse.
// Density pass (chunks of kDensityChunk): eighths 210 us, launch order 203 us and 176 MiB
// read, chunks of 16 202 us and 107 MiB.  Force pass (chunks of kForceChunk), 1M fill
// after 25 steps: eighths 117.2 us, chunks of 32 114.1 us, of 128 113.8 us.
constexpr uint32_t kDensityChunk = 16;
constexpr uint32_t kForceChunk = 128;
template <uint32_t kChunk>
__device__ __forceinline__ uint32_t xcd_chunked() {
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    constexpr uint32_t kSpan = 8u * kChunk;
    if (b >= nb / kSpan * kSpan) return b;
    const uint32_t i = b / 8u;
    return ((i / kChunk) * 8u + b % 8u) * kChunk + i % kChunk;
}

// Occupancy floors.  The single-scene density pass at 8 waves per SIMD (64 VGPRs, 2
// spilled; its 128-lane blocks' list staging then fills the CU's 160 KB of LDS exactly)
// runs 5 % faster than at the 7 its 66 VGPRs allow.  The batch instances (scene lookup per
// lane) would take 79 and 130 VGPRs, one wave per SIMD below the single-scene ones; held
// at 7 and 4 waves (3 spilled registers in the batch density pass) the batched step runs
// ~1.5 % faster.
// kF: the tolerance path's approximate square root and reciprocal (pair_sqrt / pair_div).
template <class Scenes, int kU, int kB, bool kL, bool kF>
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(std::is_same_v<Scenes, OneScene> ? 8 : 7)))
void k_density(Scenes sc, const ChainBuffers b, uint32_t n) {
    __shared__ uint32_t s_list[kL ? kListLds * kB : 1];
    const uint32_t j = xcd_chunked<kDensityChunk>() * kB + threadIdx.x;
    if (j >= n) return;
    const SceneConst c = sc.scene(j);
    const float4 s = b.pred_pos[j];
    if constexpr (!kL) {
        const DensityCellsOf<GridRuns> cells{grid_runs(c, b), sample_axis(s.x, c.minx, c.maxx, c.R),
                                             sample_axis(s.y, c.miny, c.maxy, c.R),
                                             sample_axis(s.z, c.minz, c.maxz, c.R)};
        b.dens[j] = density_walk_rows<kU, kF>(c, s, b.pred_pos, cells);
        return;
    }
    uint32_t count, found;
    const float2 d = density_with_list<kU, kF>(c, j, s, b.pred_pos, grid_runs(c, b), s_list + threadIdx.x, kB,
                                           b.nlist, n, b.list_cap, count, found);
    // The longest list above half the allocated capacity, for the host to size the lists
    // (grow past a list that did not fit, halve when none comes near): the maximum is taken
    // over those lanes only (all active), one atomic per wave that has any and none once a
    // longer list is on record.  The host clears the record each time it reads it.
    const uint64_t longer = __ballot(found > (b.list_cap >> 1));
    if (longer) {
        uint32_t longest = 0;
        for (uint64_t m = longer; m; m &= m - 1ull)
            longest = max(longest, (uint32_t)__builtin_amdgcn_readlane((int)found, __ffsll((long long)m) - 1));
        if ((threadIdx.x & 63u) == (uint32_t)(__ffsll((long long)longer) - 1) &&
            __hip_atomic_load(&b.counters[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < longest)
            atomicMax(&b.counters[4], longest);
    }
    b.dens[j] = d;
    // The densities also ride in the w lanes of the particle's predicted velocity (density)
    // and predicted position (near density), so that the force pass gathers a listed
    // neighbour with two loads.  Every reader of these arrays uses their xyz only.
    reinterpret_cast<float *>(b.vel_sorted + j)[3] = d.x;
    reinterpret_cast<float *>(b.pred_pos + j)[3] = d.y;
    b.ncount[j] = (uint8_t)count;
    // entry-major copy-out: every store of the loop writes one coalesced row
    const uint32_t m = count == kListInvalid ? 0u : min(count, (uint32_t)kListLds);
    for (uint32_t k = 0; __ballot(k < m); ++k)
        if (k < m) b.nlist[(size_t)k * n + j] = s_list[k * kB + threadIdx.x];
}

// ---------------------------------------------------------------------------------------
// Force + integrate: applyPressureAndViscosityForce.comp:275-337 followed by
// updateVelocities.comp:71-119.  Neighbour velocities come from vel_sorted, which no lane
// writes (the reference races on that buffer; here every read sees the pre-force value).
// Works in cell order; the new state of cell-ordered slot j goes to hash slot hslot[j].
template <class Scenes, int kU, int kB, bool kL, bool kF>
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(4))) void k_force(Scenes sc, const ChainBuffers b,
                                                                                        uint32_t n) {
    __shared__ uint32_t s_queue[kQueue * kB];
    const uint32_t j = xcd_chunked<kForceChunk>() * kB + threadIdx.x;
    if (j >= n) return;
    const SceneConst c = sc.scene(j);
    const float4 s = b.pred_pos[j];
    const float4 P = b.pos_sorted[j];
    const float4 V = b.vel_sorted[j];
    const float2 myd = b.dens[j];
    const ForceSelf me = force_self(c, j, s, P, V, myd);
    float3 Fp = make_float3(0.0f, 0.0f, 0.0f), Fv = make_float3(0.0f, 0.0f, 0.0f);
    const uint32_t cnt = kL ? b.ncount[j] : kListInvalid;
    if (cnt != kListInvalid) {
        // the recorded neighbours, in the walk's order
        force_list<8, kF>(c, me, cnt, b.nlist, n, b.pred_pos, b.vel_sorted, Fp, Fv);
    } else {
        const ForceCellsOf<GridRuns> cells{grid_runs(c, b), force_cell(c, s)};
        force_walk_rows<kU, kF>(c, me, b.pred_pos, b.vel_sorted, b.dens, cells, s_queue + threadIdx.x, kB, Fp, Fv);
    }
    float4 po, vo, vf;
    force_finish(c, P, V, myd.x, Fp, Fv, po, vo, vf);
    const uint32_t o = b.hslot[j];
    if (b.dbg_vel_force) {
        // V.w holds the density when lists are recorded; the placement kept the input
        // velocity's own w in the capture buffer
        if (kL) vf.w = b.dbg_vel_force[j].w + 1.0f;
        b.dbg_vel_force[j] = vf;
    }
    b.pos[o] = po;
    b.vel[o] = vo;
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1u) / kBlock; }

}  // namespace

template <class Scenes>
hipError_t launch_hash(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                       uint32_t table_total, uint32_t passes, uint32_t lookback_words,
                       uint32_t bucket_total, uint32_t digit_bits) {
    // 4 keys per lane of 1024-lane hashing workgroups: enough waves to fill the chip, few
    // global histogram atomics (every workgroup adds to the same bins, and those additions
    // are serialised; each lane's keys come one after another): at 1M keys, 512 workgroups
    // of 256 lanes over 2K keys each took 20 us, 256 of 1024 lanes over 4K 14 us, 128 over
    // 8K 20 us.  The clearing workgroups run beside them.
    const uint32_t hash_blocks = (n + kHashKeys - 1u) / kHashKeys;
    const uint32_t btiles = (bucket_total + kBucketTile - 1u) / kBucketTile;
    uint32_t clear_work = table_total > lookback_words ? table_total : lookback_words;
    if (bucket_total > clear_work) clear_work = bucket_total;
    uint32_t clear_blocks = (clear_work + 4u * kHashBlock - 1u) / (4u * kHashBlock);
    if (clear_blocks < 1u) clear_blocks = 1u;
    if (clear_blocks > 512u) clear_blocks = 512u;
    auto kernel = digit_bits == 12 ? k_hash<Scenes, 12> : digit_bits == 11 ? k_hash<Scenes, 11> : k_hash<Scenes, 8>;
    hipLaunchKernelGGL(kernel, dim3(hash_blocks + clear_blocks), dim3(kHashBlock), 0, stream,
                       sc, b.pos, b.keys, n, passes, b.ghist, hash_blocks, b.table,
                       bucket_total ? b.crun : nullptr, table_total,
                       b.lookback, lookback_words, b.tile_ctr, b.counters, b.bcount, bucket_total,
                       b.blookback, btiles);
    return hipGetLastError();
}

template <class Scenes>
hipError_t launch_reorder(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                          const uint32_t *sorted_keys, const uint32_t *sorted_idx, uint32_t n,
                          bool bricks, uint32_t ghist_words) {
    hipLaunchKernelGGL(k_reorder<Scenes>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, sc,
                       sorted_keys, sorted_idx, n, b.pos, b.vel, b.pos_sorted, b.vel_sorted,
                       b.pred_pos, b.dbg_pred_vel, b.table, b.ghist, b.hfull,
                       bricks ? b.brick_flag : nullptr, b.stray, b.counters, bricks ? nullptr : b.bcount,
                       b.run_info, ghist_words);
    return hipGetLastError();
}

const LaneConfig kLaneConfigs[kNumLaneConfigs] = {
    {2, 256, true, false}, {4, 256, true, false}, {2, 128, true, false}, {4, 128, true, false},
    {2, 64, true, false}, {2, 256, false, false}, {4, 256, false, false}, {2, 128, false, false},
    {4, 128, false, false}, {2, 64, false, false},
    {4, 128, true, true}, {4, 128, false, true},  // the tolerance path's variants of 3 and 8
};

template <class Scenes, int kU, int kB, bool kL, bool kF>
hipError_t launch_density_t(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n) {
    hipLaunchKernelGGL((k_density<Scenes, kU, kB, kL, kF>), dim3((n + kB - 1u) / kB), dim3(kB), 0, stream, sc, b, n);
    return hipGetLastError();
}

template <class Scenes, int kU, int kB, bool kL, bool kF>
hipError_t launch_force_t(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n) {
    hipLaunchKernelGGL((k_force<Scenes, kU, kB, kL, kF>), dim3((n + kB - 1u) / kB), dim3(kB), 0, stream, sc, b, n);
    return hipGetLastError();
}

// Dispatches configuration `cfg` (an index into kLaneConfigs) to its instantiation.  The
// tolerance variants exist for single scenes only (batches and served scenes stay exact).
template <template <class, int, int, bool, bool> class Launch, class Scenes>
hipError_t launch_cfg(int cfg, hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n) {
    switch (cfg) {
        case 0: return Launch<Scenes, 2, 256, true, false>::run(stream, sc, b, n);
        case 1: return Launch<Scenes, 4, 256, true, false>::run(stream, sc, b, n);
        case 2: return Launch<Scenes, 2, 128, true, false>::run(stream, sc, b, n);
        case 4: return Launch<Scenes, 2, 64, true, false>::run(stream, sc, b, n);
        case 5: return Launch<Scenes, 2, 256, false, false>::run(stream, sc, b, n);
        case 6: return Launch<Scenes, 4, 256, false, false>::run(stream, sc, b, n);
        case 7: return Launch<Scenes, 2, 128, false, false>::run(stream, sc, b, n);
        case 8: return Launch<Scenes, 4, 128, false, false>::run(stream, sc, b, n);
        case 9: return Launch<Scenes, 2, 64, false, false>::run(stream, sc, b, n);
        case 10:
            if constexpr (std::is_same_v<Scenes, OneScene>) return Launch<Scenes, 4, 128, true, true>::run(stream, sc, b, n);
            else return Launch<Scenes, 4, 128, true, false>::run(stream, sc, b, n);
        case 11:
            if constexpr (std::is_same_v<Scenes, OneScene>) return Launch<Scenes, 4, 128, false, true>::run(stream, sc, b, n);
            else return Launch<Scenes, 4, 128, false, false>::run(stream, sc, b, n);
        default: return Launch<Scenes, 4, 128, true, false>::run(stream, sc, b, n);
    }
}

template <class Scenes, int kU, int kB, bool kL, bool kF>
struct DensityLaunch {
    static hipError_t run(hipStream_t st, const Scenes &sc, const ChainBuffers &b, uint32_t n) {
        return launch_density_t<Scenes, kU, kB, kL, kF>(st, sc, b, n);
    }
};

template <class Scenes, int kU, int kB, bool kL, bool kF>
struct ForceLaunch {
    static hipError_t run(hipStream_t st, const Scenes &sc, const ChainBuffers &b, uint32_t n) {
        return launch_force_t<Scenes, kU, kB, kL, kF>(st, sc, b, n);
    }
};

template <class Scenes>
hipError_t launch_density(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                          int cfg) {
    return launch_cfg<DensityLaunch>(cfg, stream, sc, b, n);
}

template <class Scenes>
hipError_t launch_force(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n, int cfg) {
    return launch_cfg<ForceLaunch>(cfg, stream, sc, b, n);
}

#define SPH_INSTANTIATE(S)                                                                       \
    template hipError_t launch_hash<S>(hipStream_t, const S &, const ChainBuffers &, uint32_t,   \
                                       uint32_t, uint32_t, uint32_t, uint32_t, uint32_t);        \
    template hipError_t launch_reorder<S>(hipStream_t, const S &, const ChainBuffers &,          \
                                          const uint32_t *, const uint32_t *, uint32_t, bool,    \
                                          uint32_t);                                             \
    template hipError_t launch_density<S>(hipStream_t, const S &, const ChainBuffers &, uint32_t, int); \
    template hipError_t launch_force<S>(hipStream_t, const S &, const ChainBuffers &, uint32_t, int);

SPH_INSTANTIATE(OneScene)
SPH_INSTANTIATE(ManyScenes)

}  // namespace sph
