// Cell-ordered layout of the per-lane neighbour passes.
//
// The reference's slot order sorts particles by a hashed cell key, which scatters
// neighbouring cells over the whole array: a wave of 64 consecutive slots touches ~600
// unrelated runs in each neighbour pass and nothing it loads is reused by the next wave.
// The neighbour sums only require that each key's run is visited whole and in its own
// (hash) order, so the runs themselves may be placed anywhere.  Every run goes to the grid
// bucket of its first particle's cell (buckets numbered x-fastest), which puts the runs of
// neighbouring cells next to each other:
//
//   (k_reorder)   at each run's first slot: the run's bucket and, by an atomic add of the
//                 run length to the bucket's count, its offset in the bucket (runs sharing
//                 a bucket land in any order), stored per key;
//   k_bucket_scan exclusive scan of the bucket counts (single pass, decoupled look-back);
//   k_place       one lane per hash slot: its cell-ordered slot, the gather of the state
//                 with the prediction (predictPositions.comp) and the hash slot of each
//                 cell-ordered slot; each run's first lane stores the run's cell-ordered
//                 (start, end) for its key (keys without particles keep the -1 the hash
//                 pass wrote);
//   k_grid        each scene's halo grid: the cell-ordered run of the key of every cell
//                 of the box grid grown by one cell per side, which the neighbour walks
//                 read instead of hashing the cell.
//
// Placement never changes a result: every particle's sums visit the same runs in the
// same order whatever the layout, so only the memory order (and speed) depends on the
// order in which the atomics were served.
#include "sph_device.hpp"
#include "sph_kernels.hpp"
#include "sph_tile_owner.hpp"

namespace sph {

namespace {

constexpr int kBlock = 256;
constexpr int kScanItems = (int)kBucketTile / kBlock;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of bcount[0, count) in place.  Tiles are taken in launch order from an
// atomic counter, so every tile's predecessors are running or done when it looks back.
__global__ __launch_bounds__(kBlock) void k_bucket_scan(uint32_t *__restrict__ bcount, uint32_t count,
                                                       uint32_t *lookback, uint32_t *tile_counter) {
    __shared__ uint32_t s_tile, s_excl;
    __shared__ uint32_t s_wave[kBlock / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kBucketTile + (uint32_t)tid * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < count ? bcount[base + k] : 0u;
        sum += v[k];
    }
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        if (i < w) wbase += s_wave[i];
        total += s_wave[i];
    }
    if (w == 0) {
        // Decoupled look-back by the first wave: lane l reads tile - 1 - l, so 64
        // predecessors are examined per round; the nearest one with an inclusive prefix
        // ends the walk (tiles before tile 0 count as a prefix of 0).
        uint32_t excl = 0;
        if (tile == 0) {
            if (lane == 0) store_relaxed(lookback, kFlagPrefix | total);
        } else {
            if (lane == 0) store_relaxed(lookback + tile, kFlagAggregate | total);
            int32_t top = (int32_t)tile - 1;
            while (true) {
                const int32_t t = top - lane;
                const uint32_t lv = t >= 0 ? load_relaxed(lookback + t) : kFlagPrefix;
                const uint32_t f = lv & kFlagMask;
                const uint64_t not_ready = __ballot(f == 0u);
                const uint64_t prefix = __ballot(f == kFlagPrefix);
                const int pl = prefix ? __ffsll((long long)prefix) - 1 : 64;
                const uint64_t need = pl >= 63 ? ~0ull : ((1ull << (pl + 1)) - 1ull);
                if (not_ready & need) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t v2 = lane <= pl ? (lv & kValueMask) : 0u;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v2 += __shfl_xor(v2, o, 64);
                excl += v2;
                if (prefix) break;
                top -= 64;
            }
            if (lane == 0) store_relaxed(lookback + tile, kFlagPrefix | (excl + total));
        }
        if (lane == 0) s_excl = excl;
    }
    __syncthreads();
    uint32_t run = s_excl + wbase + x - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < count) bcount[base + k] = run;
        run += v[k];
    }
}

// Gathers every particle into its cell-ordered slot; the first particle of each run also
// records the run's cell-ordered (start, end) for its key.
template <class Scenes>
__global__ __launch_bounds__(kBlock) void k_place(
    Scenes sc, const uint32_t *__restrict__ skeys, const uint32_t *__restrict__ sidx, uint32_t n,
    const float4 *__restrict__ pos, const float4 *__restrict__ vel, const int2 *__restrict__ table,
    const uint32_t *__restrict__ bstart, const uint2 *__restrict__ run_info,
    float4 *__restrict__ pos_sorted, float4 *__restrict__ vel_sorted, float4 *__restrict__ pred_pos,
    float4 *__restrict__ dbg_pred_vel, float4 *__restrict__ dbg_vel_force, uint32_t *__restrict__ hslot,
    int2 *__restrict__ crun, uint2 *__restrict__ towner, uint32_t *__restrict__ stray,
    uint32_t *__restrict__ counters) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const uint32_t key = skeys[j];
    const int2 t = table[key];
    const uint32_t head = (uint32_t)t.x;
    const uint2 ri = run_info[key];
    const uint32_t j2 = bstart[ri.x] + ri.y + (j - head);
    if (j == head) crun[key] = make_int2((int)j2, (int)j2 + (t.y - t.x));
    const uint32_t idx = sidx ? sidx[j] : j;
    const float4 P = pos[idx];
    const float4 V = vel[idx];
    const SceneConst c = sc.scene(j);
    float3 p = make_float3(P.x, P.y, P.z), v = make_float3(V.x, V.y, V.z);
    kick_drift<false>(c, p, v);
    pos_sorted[j2] = P;
    vel_sorted[j2] = V;
    pred_pos[j2] = make_float4(p.x, p.y, p.z, 1.0f);
    if (dbg_pred_vel) dbg_pred_vel[j2] = make_float4(v.x, v.y, v.z, 1.0f);
    // the input velocity's w for the debug capture of the force pass's velocities[j] + (.., 1):
    // the list-recording density pass reuses vel_sorted[].w for the density
    if (dbg_vel_force) dbg_vel_force[j2] = make_float4(0.0f, 0.0f, 0.0f, V.w);
    hslot[j2] = j;
    if (towner) {
        const uint2 own = tile_owner(c, make_float3(P.x, P.y, P.z), p);
        towner[j2] = own;
        // strays are listed for the per-lane code of the tiled launches; one atomic per wave
        const bool is_stray = own.x == kNoTile;
        const uint64_t m = __ballot(is_stray);
        if (m) {
            const int lane = threadIdx.x & 63;
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&counters[1], (uint32_t)__popcll(m));
            base = __shfl(base, leader, 64);
            if (is_stray) stray[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = j2;
        }
    }
}

// The halo grid: each cell's key's cell-ordered run (one lookup per cell).
template <class Scenes>
__global__ __launch_bounds__(kBlock) void k_grid(Scenes sc, uint32_t grid_total,
                                                 const int2 *__restrict__ crun, int2 *__restrict__ grid) {
    const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t g0 = __builtin_amdgcn_readfirstlane(g);  // the wave's first entry
    if (g >= grid_total) return;
    const SceneConst c = sc.scene_of_grid_near(g0, g);
    const uint32_t hx = (uint32_t)c.gx + 2u, hy = (uint32_t)c.gy + 2u;
    const uint32_t l = g - c.grid_base;
    const uint32_t lx = l % hx, ly = (l / hx) % hy, lz = l / (hx * hy);
    // cell (lx-1, ly-1, lz-1); -1 wraps to 2^32-1 as in the reference's u32 hash
    grid[g] = crun[c.table_base + (cell_hash(lx - 1u, ly - 1u, lz - 1u) & c.mask)];
}

inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1u) / kBlock; }

}  // namespace

hipError_t launch_bucket_scan(hipStream_t stream, const ChainBuffers &b, uint32_t bucket_total) {
    const uint32_t tiles = (bucket_total + kBucketTile - 1u) / kBucketTile;
    hipLaunchKernelGGL(k_bucket_scan, dim3(tiles), dim3(kBlock), 0, stream, b.bcount, bucket_total,
                       b.blookback, b.tile_ctr + 4);
    return hipGetLastError();
}

template <class Scenes>
hipError_t launch_place(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                        const uint32_t *sorted_keys, const uint32_t *sorted_idx, uint32_t n,
                        uint32_t grid_total, bool tiles) {
    hipLaunchKernelGGL(k_place<Scenes>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, sc, sorted_keys,
                       sorted_idx, n, b.pos, b.vel, b.table, b.bcount, b.run_info, b.pos_sorted,
                       b.vel_sorted, b.pred_pos, b.dbg_pred_vel, b.dbg_vel_force, b.hslot, b.crun,
                       tiles ? b.towner : nullptr, b.stray, b.counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || grid_total == 0) return e;
    hipLaunchKernelGGL(k_grid<Scenes>, dim3(blocks_for(grid_total)), dim3(kBlock), 0, stream, sc,
                       grid_total, b.crun, b.grid);
    return hipGetLastError();
}

#define SPH_LAYOUT_INSTANTIATE(S)                                                                \
    template hipError_t launch_place<S>(hipStream_t, const S &, const ChainBuffers &,            \
                                        const uint32_t *, const uint32_t *, uint32_t, uint32_t, bool);

SPH_LAYOUT_INSTANTIATE(OneScene)
SPH_LAYOUT_INSTANTIATE(ManyScenes)

}  // namespace sph
