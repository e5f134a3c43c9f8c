// Stable LSD radix sort of (cell key, particle index) pairs for gfx950.
//
// Replaces multiRadixSortHistograms.comp + multiRadixSort.comp and their 4-pass driver
// (commands.c:106-161).  The reference's sort is stable (WG-ordered global offsets
// multiRadixSort.comp:57-89, in-WG ordered scatter :96-139), so its output is the pairs
// ordered by (key, index); this sort produces exactly that order.
//
// Design (one kernel per 8-bit digit, "onesweep" style):
//   * the digit histograms of every pass are accumulated once, by the hash kernel;
//   * each workgroup takes the next tile in launch order from an atomic counter, ranks its
//     keys with wave64 ballots (8 ballots give each lane the mask of lanes holding the same
//     digit; popcount below the lane is its rank), combines the per-wave counters in LDS,
//     publishes its per-digit counts and resolves its global offsets by decoupled
//     look-back over earlier tiles, then scatters through LDS so that each digit's run is
//     written contiguously.
//   * Only ceil(log2(T)/8) passes run: keys are below the table size T.
#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

// Exclusive prefix sum over the block's 256 threads (one value each).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_wave_tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave_tot[w] = x;
    __syncthreads();
    uint32_t base = 0;
    for (int i = 0; i < w; ++i) base += s_wave_tot[i];
    __syncthreads();
    return base + x - v;
}

__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int kItems>
__global__ __launch_bounds__(kBlock) void k_radix_pass(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n, uint32_t shift,
    uint32_t bits, const uint32_t *__restrict__ ghist, uint32_t *lookback,
    uint32_t *tile_counter) {
    constexpr int kTile = kBlock * kItems;
    __shared__ uint32_t s_whist[kWaves][256];
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ uint32_t s_digit_off[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_tot[kWaves];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < kWaves * 256; i += kBlock) (&s_whist[0][0])[i] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t wave_base = tile * kTile + (uint32_t)w * (64u * kItems);
    const uint32_t dmask = (1u << bits) - 1u;
    const uint64_t lt_mask = (1ull << lane) - 1ull;

    uint32_t key[kItems], val[kItems], rank[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t pos = wave_base + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = pos < n;
        key[k] = valid ? kin[pos] : 0u;
        val[k] = valid ? (vin ? vin[pos] : pos) : 0u;
    }

    // Rank within the wave, item by item, so equal digits keep their input order.
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t pos = wave_base + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = pos < n;
        const uint32_t d = (key[k] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
        for (uint32_t b = 0; b < bits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
            const uint32_t prior = s_whist[w][d];
            r = prior + below;
            if (below == 0u) s_whist[w][d] = prior + (uint32_t)__popcll(peers);
        }
        rank[k] = r;
    }
    __syncthreads();

    // Thread `tid` owns digit `tid`: per-wave exclusive offsets and the tile's count.
    uint32_t tcount = 0;
    {
        uint32_t sum = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) {
            const uint32_t c = s_whist[ww][tid];
            s_whist[ww][tid] = sum;
            sum += c;
        }
        tcount = sum;
    }
    const uint32_t gexcl = block_exclusive_scan(ghist[tid], s_tot);
    const uint32_t local_off = block_exclusive_scan(tcount, s_tot);
    s_digit_off[tid] = local_off;

    // Decoupled look-back: the global start of this tile's digit-`tid` run.
    uint32_t *mine = lookback + (size_t)tile * 256u + (uint32_t)tid;
    uint32_t excl;
    if (tile == 0) {
        excl = gexcl;
        store_relaxed(mine, kFlagPrefix | (excl + tcount));
    } else {
        store_relaxed(mine, kFlagAggregate | tcount);
        uint32_t acc = 0;
        int32_t t = (int32_t)tile - 1;
        while (true) {
            const uint32_t v = load_relaxed(lookback + (size_t)t * 256u + (uint32_t)tid);
            const uint32_t f = v & kFlagMask;
            if (f == 0u) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            acc += v & kValueMask;
            if (f == kFlagPrefix) break;
            --t;
        }
        excl = acc;
        store_relaxed(mine, kFlagPrefix | (excl + tcount));
    }
    s_gbase[tid] = excl;
    __syncthreads();

    // Order the tile by digit in LDS, then write each digit's run contiguously.
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t pos = wave_base + (uint32_t)k * 64u + (uint32_t)lane;
        if (pos < n) {
            const uint32_t d = (key[k] >> shift) & dmask;
            const uint32_t lp = s_digit_off[d] + s_whist[w][d] + rank[k];
            s_keys[lp] = key[k];
            s_vals[lp] = val[k];
        }
    }
    __syncthreads();
    const uint32_t tile_begin = tile * kTile;
    const uint32_t tile_n = min((uint32_t)kTile, n - tile_begin);
    for (uint32_t i = tid; i < tile_n; i += kBlock) {
        const uint32_t kk = s_keys[i];
        const uint32_t d = (kk >> shift) & dmask;
        const uint32_t gp = s_gbase[d] + (i - s_digit_off[d]);
        kout[gp] = kk;
        vout[gp] = s_vals[i];
    }
}

constexpr int kItemsDefault = 16;  // 4096-key tiles: half the look-back chain of 2048-key tiles

}  // namespace

uint32_t radix_tile_size() { return (uint32_t)(kBlock * kItemsDefault); }

hipError_t radix_pass(hipStream_t stream, const uint32_t *kin, const uint32_t *vin,
                      uint32_t *kout, uint32_t *vout, uint32_t n, uint32_t shift,
                      uint32_t bits, const uint32_t *ghist, uint32_t *lookback,
                      uint32_t *tile_counter) {
    const uint32_t tiles = (n + radix_tile_size() - 1u) / radix_tile_size();
    hipLaunchKernelGGL(k_radix_pass<kItemsDefault>, dim3(tiles), dim3(kBlock), 0, stream, kin,
                       vin, kout, vout, n, shift, bits, ghist, lookback, tile_counter);
    return hipGetLastError();
}

}  // namespace sph
