// Allocation-free substep chain over caller-owned device state, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Every intermediate buffer of the per-lane chain lives in one
// file-scope device array; the kernels are the ones every world runs, so results are
// bit-identical to sph_world_step on the same state and parameters.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/sph_world.h"
#include "sph_host.hpp"
#include "sph_kernels.hpp"

namespace sph {

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Bytes of the intermediates of one scene of n particles and T table cells on the per-lane
// path without force lists: keys, two radix ping-pong pairs, predicted / sorted positions,
// sorted velocities, densities, hash slots (80 B per particle); start/end, run placement and
// cell-ordered run per table cell (24 B); digit histograms, look-back words and counters;
// bucket counts (at most 2n + 1024 buckets) and their look-back words; the halo grid (at
// most 16n + 4096 cells).
constexpr uint64_t scratch_need(uint64_t n, uint64_t T, uint64_t passes, uint64_t buckets, uint64_t grid) {
    const uint64_t tiles = (n + 4095) / 4096;
    return align_up(4 * n) * 5 + align_up(16 * n) * 3 + align_up(8 * n) + align_up(4 * n) +
           align_up(8 * T) * 3 + align_up(4 * kGhistCopies * 256 * 4) + align_up(4 * passes * tiles * 256) +
           align_up(4 * kTileCounters) + align_up(4 * 8) + align_up(4 * buckets) +
           align_up(4 * ((buckets + kBucketTile - 1) / kBucketTile)) + align_up(8 * grid);
}

constexpr uint64_t kMaxN = SPH_NOALLOC_MAX_PARTICLES;
constexpr uint64_t kMaxT = SPH_NOALLOC_MAX_TABLE_CELLS;
constexpr uint64_t kScratchBytes = scratch_need(kMaxN, kMaxT, 3, 2 * kMaxN + 1024, 16 * kMaxN + 4096);
static_assert(kScratchBytes <= (72ull << 20), "allocation-free scratch exceeds its stated 72 MiB");

// The scratch array.  Its name is a host-side stand-in in host code: the device address of
// each GPU's copy comes from hipGetSymbolAddress (scratch_on_current_device).
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(SPH_ERR_DEVICE, "device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SPH_OK;
}

__global__ void k_zero_u32(uint32_t *p, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// Carves the chain's buffers out of the scratch in a fixed order.
struct Carver {
    unsigned char *base;
    uint64_t used = 0;
    template <class T>
    T *take(uint64_t bytes) {
        T *p = reinterpret_cast<T *>(base + used);
        used += align_up(bytes);
        return p;
    }
};

int is_device_memory(const void *p) {
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

}  // namespace

}  // namespace sph

extern "C" int sph_step_noalloc(const sph_params *params, uint32_t n, float *pos4_dev, float *vel4_dev,
                                uint32_t frames, void *stream) {
    using namespace sph;
    int rc = validate_params(params);
    if (rc) return rc;
    if (n == 0) return fail(SPH_ERR_INVALID_ARG, "particle count must be > 0");
    if (n > kMaxN) return fail(SPH_ERR_CAPACITY, "%u particles exceed the allocation-free maximum of %llu", n,
                               (unsigned long long)kMaxN);
    const uint32_t T = params->table_cells ? params->table_cells : n;
    if (T > kMaxT) return fail(SPH_ERR_CAPACITY, "%u table cells exceed the allocation-free maximum of %llu", T,
                               (unsigned long long)kMaxT);
    if (!pos4_dev || !vel4_dev) return fail(SPH_ERR_INVALID_ARG, "state pointer is NULL");
    if (pos4_dev == vel4_dev) return fail(SPH_ERR_INVALID_ARG, "positions and velocities alias");
    if (((uintptr_t)pos4_dev | (uintptr_t)vel4_dev) & 15u)
        return fail(SPH_ERR_INVALID_ARG, "state buffers must be 16-byte aligned");
    if (!is_device_memory(pos4_dev) || !is_device_memory(vel4_dev))
        return fail(SPH_ERR_INVALID_ARG, "state buffers must be device memory");
    const uint64_t substeps = (uint64_t)frames * params->substeps;
    if (substeps == 0) return SPH_OK;

    void *base = nullptr;
    rc = scratch_on_current_device(&base);
    if (rc) return rc;

    OneScene scene{};
    scene.c = make_scene_const(*params, T, 0, 0, n, 0, 0, 0);
    Chain chain;
    chain.configure(n, T, 0, scene_buckets(scene.c), scene.c.grid_cells);
    chain.bricks = false;
    chain.lane_cfg = 8;  // 4 candidates per scan step, 128-lane workgroups, force cell walk
    const uint64_t need = scratch_need(n, T, chain.passes, chain.bucket_total, chain.grid_total);
    if (need > kScratchBytes)
        return fail(SPH_ERR_CAPACITY, "scene needs %llu scratch bytes, %llu available", (unsigned long long)need,
                    (unsigned long long)kScratchBytes);

    Carver cv{static_cast<unsigned char *>(base)};
    ChainBuffers &b = chain.b;
    const uint64_t N = n;
    b.keys = cv.take<uint32_t>(4 * N);
    b.sort_k[0] = cv.take<uint32_t>(4 * N);
    b.sort_v[0] = cv.take<uint32_t>(4 * N);
    b.sort_k[1] = cv.take<uint32_t>(4 * N);
    b.sort_v[1] = cv.take<uint32_t>(4 * N);
    b.pred_pos = cv.take<float4>(16 * N);
    b.pos_sorted = cv.take<float4>(16 * N);
    b.vel_sorted = cv.take<float4>(16 * N);
    b.dens = cv.take<float2>(8 * N);
    b.hslot = cv.take<uint32_t>(4 * N);
    b.table = cv.take<int2>(8ull * T);
    b.run_info = cv.take<uint2>(8ull * T);
    b.crun = cv.take<int2>(8ull * T);
    b.ghist = cv.take<uint32_t>(4 * kGhistCopies * 256 * 4);
    const uint64_t tiles = (N + 4095) / 4096;
    b.lookback = cv.take<uint32_t>(4 * (uint64_t)chain.passes * tiles * 256);
    b.tile_ctr = cv.take<uint32_t>(4 * kTileCounters);
    b.counters = cv.take<uint32_t>(4 * 8);
    b.bcount = cv.take<uint32_t>(4ull * chain.bucket_total);
    b.blookback = cv.take<uint32_t>(4ull * ((chain.bucket_total + kBucketTile - 1) / kBucketTile));
    b.grid = cv.take<int2>(8ull * chain.grid_total);
    b.pos = reinterpret_cast<float4 *>(pos4_dev);
    b.vel = reinterpret_cast<float4 *>(vel4_dev);
    if (radix_tile_size() != 4096u || cv.used != need)
        return fail(SPH_ERR_STATE, "allocation-free scratch layout is inconsistent");

    hipStream_t s = static_cast<hipStream_t>(stream);
    // The array keeps the previous call's contents: the digit histograms must start at zero
    // (each substep's reorder clears them for the next); a kernel, not a memset, so that a
    // caller may capture this call into a graph.  Everything else is cleared by the chain.
    hipLaunchKernelGGL(k_zero_u32, dim3(kGhistCopies * 4u), dim3(256), 0, s, b.ghist, kGhistCopies * 4u * 256u);
    hipError_t e = hipGetLastError();
    for (uint64_t k = 0; k < substeps && e == hipSuccess; ++k) e = chain.enqueue(s, scene, nullptr);
    return e == hipSuccess ? SPH_OK : hip_fail(e, "allocation-free substep launch");
}
