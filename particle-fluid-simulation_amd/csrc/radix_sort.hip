// Stable LSD radix sort of (cell key, particle index) pairs for gfx950.
//
// Replaces multiRadixSortHistograms.comp + multiRadixSort.comp and their 4-pass driver
// (commands.c:106-161).  The reference's sort is stable (WG-ordered global offsets
// multiRadixSort.comp:57-89, in-WG ordered scatter :96-139), so its output is the pairs
// ordered by (key, index); this sort produces exactly that order.
//
// Design (one kernel per digit of 8, 11 or 12 bits, chosen per size bucket; "onesweep" style):
//   * the digit histograms of every pass are accumulated once, by the hash kernel;
//   * each workgroup takes the next tile in launch order from an atomic counter, ranks its
//     keys with wave64 ballots (8 ballots give each lane the mask of lanes holding the same
//     digit; popcount below the lane is its rank), combines the per-wave counters in LDS,
//     publishes its per-digit counts and resolves its global offsets by decoupled
//     look-back over earlier tiles, then scatters through LDS so that each digit's run is
//     written contiguously.
//   * Only ceil(log2(T)/w) passes of w-bit digits run: keys are below the table size T.
#include <type_traits>

#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

namespace {

constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr uint32_t kTileKeys = 4096;  // keys per tile, every digit width

// Exclusive prefix sum over the block's kBlk threads (one value each).
template <int kBlk>
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

// One LSD pass over kW-bit digits (kW = 8, 11 or 12: 256, 2048 or 4096 bins) with kBlk
// threads per 4096-key tile.  In the per-digit phase (tile counts, scans, look-back) the
// first kBins / kDPT threads each own kDPT consecutive digits.  A tile's latency, not
// bandwidth, bounds the pass: more waves per tile hide the loads, the ballots and the LDS
// round trips of the ranking.
template <int kBlk, int kW>
__global__ __launch_bounds__(kBlk) void k_radix_pass(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n, uint32_t shift,
    uint32_t bits, const uint32_t *__restrict__ ghist, uint32_t ghist_stride, uint32_t *lookback,
    uint32_t *tile_counter) {
    constexpr int kWaves = kBlk / 64;
    constexpr int kItems = (int)kTileKeys / kBlk;
    constexpr int kTile = (int)kTileKeys;
    constexpr int kBins = 1 << kW;
    constexpr int kDPT = kBins >= kBlk ? kBins / kBlk : 1;
    constexpr int kOwners = kBins / kDPT;
    static_assert(kItems * kBlk == kTile, "tile must split evenly over the block");
    // Per wave and digit: the count, then the LDS start of the wave's run (both at most
    // kTile), in 16-bit words where 32-bit ones would not fit LDS beside the tile.
    using WaveCount = std::conditional_t<(kWaves * kBins * 4 > 64 * 1024), uint16_t, uint32_t>;
    static_assert(kTile < 65536, "16-bit wave counts hold every LDS position of a tile");
    __shared__ WaveCount s_whist[kWaves][kBins];
    __shared__ uint32_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];
    __shared__ int32_t s_delta[kBins];           // global start minus LDS start of each digit
    __shared__ uint32_t s_tot[kWaves];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const bool owner = tid < kOwners;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < kWaves * kBins; i += kBlk) (&s_whist[0][0])[i] = 0u;
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

    // Rank within the wave, item by item, so equal digits keep their input order.  All kW
    // digit bits are compared: bits at and above `bits` are zero in every lane and leave
    // the mask unchanged.
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t pos = wave_base + (uint32_t)k * 64u + (uint32_t)lane;
        const bool valid = pos < n;
        const uint32_t d = (key[k] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kW; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
            const uint32_t prior = s_whist[w][d];
            r = prior + below;
            if (below == 0u) s_whist[w][d] = (WaveCount)(prior + (uint32_t)__popcll(peers));
        }
        rank[k] = r;
    }
    __syncthreads();

    // Per owned digit: the tile's count, with each wave's count turned into the exclusive
    // count of the waves before it.
    uint32_t tcount[kDPT], gcount[kDPT];
    uint32_t tsum = 0, gsum = 0;
    if (owner) {
#pragma unroll
        for (int q = 0; q < kDPT; ++q) {
            const int d = tid * kDPT + q;
            uint32_t sum = 0;
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) {
                const uint32_t c = s_whist[ww][d];
                s_whist[ww][d] = (WaveCount)sum;
                sum += c;
            }
            tcount[q] = sum;
            tsum += sum;
            uint32_t gc = 0;
#pragma unroll
            for (uint32_t r = 0; r < kGhistCopies; ++r) gc += ghist[r * ghist_stride + (uint32_t)d];
            gcount[q] = gc;
            gsum += gc;
        }
    }
    const uint32_t gbase = block_exclusive_scan<kBlk>(gsum, s_tot);
    const uint32_t lbase = block_exclusive_scan<kBlk>(tsum, s_tot);

    // Decoupled look-back for the owned digits, walked together: the global start of this
    // tile's run of each digit.
    uint32_t excl[kDPT], acc[kDPT], loff[kDPT];
    int32_t at[kDPT];
    uint32_t pending = 0;
    if (owner) {
        uint32_t g = gbase, l = lbase;
#pragma unroll
        for (int q = 0; q < kDPT; ++q) {
            const int d = tid * kDPT + q;
            loff[q] = l;
            l += tcount[q];
            const uint32_t gd = g;
            g += gcount[q];
            uint32_t *mine = lookback + (size_t)tile * kBins + (uint32_t)d;
            if (tile == 0) {
                excl[q] = gd;
                store_relaxed(mine, kFlagPrefix | (gd + tcount[q]));
            } else {
                store_relaxed(mine, kFlagAggregate | tcount[q]);
                acc[q] = 0u;
                at[q] = (int32_t)tile - 1;
                pending |= 1u << q;
            }
        }
    }
    // Each step reads kLook predecessors of a digit at once (nearest first) and sums their
    // counts up to the first inclusive prefix, so that prefixes spread up to kLook tiles per
    // memory round trip.  A predecessor that has not published yet ends the step there; the
    // next step resumes at it.  Tile 0 always publishes a prefix, so no step passes it.
    // Measured on the 1M-key sort: 4 reads beat 1, 2 and 8; 16 and more congest the L2.
    constexpr int kLook = kDPT <= 2 ? 4 : 1;
    while (pending) {
        bool waited = false;
#pragma unroll
        for (int q = 0; q < kDPT; ++q) {
            if (!((pending >> q) & 1u)) continue;
            const int d = tid * kDPT + q;
            uint32_t v[kLook];
#pragma unroll
            for (int k = 0; k < kLook; ++k)
                v[k] = at[q] - k >= 0 ? load_relaxed(lookback + (size_t)(at[q] - k) * kBins + (uint32_t)d) : 0u;
            bool done = false, stalled = false;
#pragma unroll
            for (int k = 0; k < kLook; ++k) {
                if (done || stalled) break;
                const uint32_t f = v[k] & kFlagMask;
                if (f == 0u) {
                    stalled = true;
                    waited = true;
                    at[q] -= k;
                    break;
                }
                acc[q] += v[k] & kValueMask;
                if (f == kFlagPrefix) done = true;
            }
            if (done) {
                excl[q] = acc[q];
                store_relaxed(lookback + (size_t)tile * kBins + (uint32_t)d, kFlagPrefix | (acc[q] + tcount[q]));
                pending &= ~(1u << q);
            } else if (!stalled) {
                at[q] -= kLook;
            }
        }
        if (waited) __builtin_amdgcn_s_sleep(1);
    }
    if (owner) {
#pragma unroll
        for (int q = 0; q < kDPT; ++q) {
            const int d = tid * kDPT + q;
            s_delta[d] = (int32_t)excl[q] - (int32_t)loff[q];
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) s_whist[ww][d] = (WaveCount)(s_whist[ww][d] + loff[q]);
        }
    }
    __syncthreads();

    // Order the tile by digit in LDS, then write each digit's run contiguously.
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint32_t pos = wave_base + (uint32_t)k * 64u + (uint32_t)lane;
        if (pos < n) {
            const uint32_t d = (key[k] >> shift) & dmask;
            const uint32_t lp = s_whist[w][d] + rank[k];
            s_keys[lp] = key[k];
            s_vals[lp] = val[k];
        }
    }
    __syncthreads();
    const uint32_t tile_begin = tile * kTile;
    const uint32_t tile_n = min((uint32_t)kTile, n - tile_begin);
    for (uint32_t i = tid; i < tile_n; i += kBlk) {
        const uint32_t kk = s_keys[i];
        const uint32_t d = (kk >> shift) & dmask;
        const uint32_t gp = (uint32_t)(s_delta[d] + (int32_t)i);
        kout[gp] = kk;
        vout[gp] = s_vals[i];
    }
}

// Threads per tile for each digit width.  8- and 11-bit digits: 1024 threads (16 waves, four
// per SIMD) rank the 4096 keys with 4 items each (the 11-bit per-wave counts in 16-bit
// words: 64 KB).  The 12-bit pass keeps 256 threads: 16 waves of 4096 16-bit counts would
// not fit LDS beside the tile.
constexpr int kBlock8 = 1024;
constexpr int kBlock11 = 1024;
constexpr int kBlockWide = 256;

}  // namespace

uint32_t radix_tile_size() { return kTileKeys; }

hipError_t radix_pass(hipStream_t stream, const uint32_t *kin, const uint32_t *vin,
                      uint32_t *kout, uint32_t *vout, uint32_t n, uint32_t shift,
                      uint32_t bits, const uint32_t *ghist, uint32_t ghist_stride, uint32_t *lookback,
                      uint32_t *tile_counter, uint32_t digit_bits) {
    const uint32_t tiles = (n + kTileKeys - 1u) / kTileKeys;
    switch (digit_bits) {
        case 11:
            hipLaunchKernelGGL((k_radix_pass<kBlock11, 11>), dim3(tiles), dim3(kBlock11), 0, stream, kin,
                               vin, kout, vout, n, shift, bits, ghist, ghist_stride, lookback, tile_counter);
            break;
        case 12:
            hipLaunchKernelGGL((k_radix_pass<kBlockWide, 12>), dim3(tiles), dim3(kBlockWide), 0, stream, kin,
                               vin, kout, vout, n, shift, bits, ghist, ghist_stride, lookback, tile_counter);
            break;
        default:
            hipLaunchKernelGGL((k_radix_pass<kBlock8, 8>), dim3(tiles), dim3(kBlock8), 0, stream, kin,
                               vin, kout, vout, n, shift, bits, ghist, ghist_stride, lookback, tile_counter);
    }
    return hipGetLastError();
}

}  // namespace sph
