// Brick-tiled density and force + integrate passes for gfx950.
//
// The reference walks the 27 neighbour cells of every particle straight from the particle
// buffers (calculateDensities.comp:114-131, applyPressureAndViscosityForce.comp:288-325).
// Particles are stored in cell-key order, which is a hash order: neighbouring cells are far
// apart in memory and a wave's 64 lanes gather from ~20 unrelated places per candidate.
//
// Here a workgroup owns a brick of kBrick^3 cells.  It looks up the (kBrick+2)^3 halo cells
// of the brick once, copies their runs (predicted positions; for the force pass also
// velocities and densities) into LDS, and evaluates every particle whose current cell lies
// in the brick with the candidate loop reading LDS.  Each staged particle is reused by all
// brick particles around it instead of being gathered again from L2 for each of them.
//
// Exactness: a particle still visits its 27 cells in the reference's (dz, dy, dx) order, a
// run is the same contiguous slot range [start, end) whether it is read from LDS or from
// global memory, and the float expressions are the shared ones of sph_pairs.hpp, so the
// results are bit-identical to the one-lane-per-particle kernels.  Cells outside the staged
// halo (a particle whose predicted cell left the brick) or runs that did not fit in LDS are
// read from global memory.
//
// Work distribution: the reorder kernel marks the brick of every particle; k_queue compacts
// the marked bricks into a queue; persistent workgroups pull queue items with an atomic
// counter until it runs dry.  Particles outside every brick grid ("strays", only possible
// when the wall margin exceeds half the box) are processed one per lane after the bricks.
#include "sph_device.hpp"
#include "sph_kernels.hpp"
#include "sph_pairs.hpp"

namespace sph {

namespace {

constexpr int kBlock = 256;
constexpr int kCells = kBrick * kBrick * kBrick;
constexpr int kCapD = 1536;    // staged particles, density pass: 16 B each, 24 KiB
constexpr int kCapF = 1024;    // staged particles, force pass: 32 B each, 32 KiB
constexpr int kOwnCap = 512;   // particles evaluated per round of a brick
static_assert(kHalo <= kBlock && kCells <= kBlock, "one thread per halo cell");
static_assert(kHalo < 256, "halo cell ids are stored in 8 bits");

// Exclusive prefix sum over the 256 threads; also returns the block total.
__device__ __forceinline__ uint32_t scan256(uint32_t v, uint32_t *s_tmp, uint32_t &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        const uint32_t t = s_tmp[i];
        if (i < w) base += t;
        total += t;
    }
    __syncthreads();
    return base + x - v;
}

// Owned particles of the current round: slot and LDS index (0xFFFF = not staged).
struct Owned {
    uint32_t slot[kOwnCap];
    uint16_t lds[kOwnCap];
};

struct BrickCtx {
    int ox, oy, oz;  // cell coordinate of halo cell (0,0,0)
};

__device__ __forceinline__ BrickCtx brick_origin(const SceneConst &c, uint32_t brick) {
    const uint32_t lb = brick - c.brick_base;
    const uint32_t bx = lb % c.nbx, by = (lb / c.nbx) % c.nby, bz = lb / (c.nbx * c.nby);
    return BrickCtx{(int)bx * kBrick - 1, (int)by * kBrick - 1, (int)bz * kBrick - 1};
}

__device__ __forceinline__ bool interior(int h) {
    const int hx = h % kH1, hy = (h / kH1) % kH1, hz = h / (kH1 * kH1);
    return hx >= 1 && hx <= kBrick && hy >= 1 && hy <= kBrick && hz >= 1 && hz <= kBrick;
}

// Looks up the halo runs (one thread per halo cell) and assigns LDS space in halo order
// while it lasts; writes the owning halo cell of every staged entry into s_cellof.
__device__ __forceinline__ void halo_lookup(const SceneConst &c, const int2 *__restrict__ table,
                                            const BrickCtx &bc, int cap, Halo &H,
                                            uint8_t *s_cellof, uint32_t *s_tmp) {
    const int tid = threadIdx.x;
    uint32_t cnt = 0;
    int gst = -1;
    if (tid < kHalo) {
        const int hx = tid % kH1, hy = (tid / kH1) % kH1, hz = tid / (kH1 * kH1);
        const uint32_t x = (uint32_t)(bc.ox + hx), y = (uint32_t)(bc.oy + hy), z = (uint32_t)(bc.oz + hz);
        const uint32_t h = cell_hash(x, y, z);
        const int2 se = table[c.table_base + (h & c.mask)];
        if (se.x != -1) {
            cnt = (uint32_t)(se.y - se.x);
            gst = se.x;
        }
        H.hash[tid] = h;
    }
    uint32_t total;
    const uint32_t off = scan256(cnt, s_tmp, total);
    if (tid == 0) H.staged = 0;
    __syncthreads();
    if (tid < kHalo) {
        const bool fits = off + cnt <= (uint32_t)cap;
        H.oc[tid] = make_int2(fits ? (int)off : -1, (int)cnt);
        H.gst[tid] = gst;
        if (fits) {
            for (uint32_t k = 0; k < cnt; ++k) s_cellof[off + k] = (uint8_t)tid;
            if (cnt) atomicMax(&H.staged, off + cnt);
        }
    }
    __syncthreads();
}

// Builds the owned list of entries [chunk, chunk + kOwnCap) of the brick's particles in
// interior-cell order.  Ownership of staged entries is read from s_own_flag, of unstaged
// interior runs from the slots' full hashes.  Returns the total owned count.
__device__ __forceinline__ uint32_t owned_round(const Halo &H, const uint8_t *s_own_flag,
                                                const uint32_t *__restrict__ hfull, uint32_t chunk,
                                                Owned &O, uint32_t *s_tmp) {
    const int tid = threadIdx.x;
    int h = 0, off = -1, cnt = 0, gst = 0;
    uint32_t own = 0;
    if (tid < kCells) {
        const int lx = tid % kBrick + 1, ly = (tid / kBrick) % kBrick + 1, lz = tid / (kBrick * kBrick) + 1;
        h = lx + kH1 * (ly + kH1 * lz);
        off = H.oc[h].x;
        cnt = H.oc[h].y;
        gst = H.gst[h];
        for (int k = 0; k < cnt; ++k)
            own += (off >= 0 ? s_own_flag[off + k] != 0 : hfull[gst + k] == H.hash[h]) ? 1u : 0u;
    }
    uint32_t total;
    uint32_t idx = scan256(own, s_tmp, total);
    if (own && idx < chunk + kOwnCap && idx + own > chunk) {
        for (int k = 0; k < cnt; ++k) {
            const bool mine = off >= 0 ? s_own_flag[off + k] != 0 : hfull[gst + k] == H.hash[h];
            if (!mine) continue;
            if (idx >= chunk && idx < chunk + kOwnCap) {
                O.slot[idx - chunk] = (uint32_t)(gst + k);
                O.lds[idx - chunk] = off >= 0 ? (uint16_t)(off + k) : (uint16_t)0xFFFFu;
            }
            ++idx;
        }
    }
    __syncthreads();
    return total;
}

// Pulls the next work item of a persistent workgroup.
__device__ __forceinline__ uint32_t next_item(uint32_t *drain, uint32_t *s_item) {
    if (threadIdx.x == 0) *s_item = atomicAdd(drain, 1u);
    __syncthreads();
    const uint32_t item = *s_item;
    __syncthreads();
    return item;
}

// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_queue(uint32_t *__restrict__ flag, uint32_t total,
                                                  uint32_t *__restrict__ queue,
                                                  uint32_t *__restrict__ counters) {
    const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
    const bool f = b < total && flag[b] != 0u;
    if (f) flag[b] = 0u;
    const uint64_t m = __ballot(f);
    if (m) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counters[0], (uint32_t)__popcll(m));
        base = __shfl(base, leader, 64);
        if (f) queue[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = b;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2) counters[2 + threadIdx.x] = 0u;
}

template <class Scenes>
__global__ __launch_bounds__(kBlock) void k_density_brick(
    Scenes sc, const uint32_t *__restrict__ queue, uint32_t *__restrict__ counters,
    const float4 *__restrict__ pred_pos, const int2 *__restrict__ table,
    const uint32_t *__restrict__ hfull, const uint32_t *__restrict__ stray,
    float2 *__restrict__ dens) {
    __shared__ Halo H;
    __shared__ Owned O;
    __shared__ float4 s_pos[kCapD];
    __shared__ uint8_t s_cellof[kCapD];
    __shared__ uint8_t s_flag[kCapD];
    __shared__ uint32_t s_tmp[kBlock / 64];
    __shared__ uint32_t s_item;
    const uint32_t qn = counters[0], ns = counters[1];
    const uint32_t items = qn + (ns + kBlock - 1u) / kBlock;
    for (;;) {
        const uint32_t item = next_item(&counters[2], &s_item);
        if (item >= items) break;
        if (item >= qn) {
            const uint32_t idx = (item - qn) * kBlock + threadIdx.x;
            if (idx < ns) {
                const uint32_t j = stray[idx];
                const SceneConst c = sc.scene(j);
                const float4 s = pred_pos[j];
                const DensityCells<false> cells{CellRuns<false>{c.table_base, c.mask, table, nullptr, 0, 0, 0},
                                                sample_axis(s.x, c.minx, c.maxx, c.R),
                                                sample_axis(s.y, c.miny, c.maxy, c.R),
                                                sample_axis(s.z, c.minz, c.maxz, c.R)};
                dens[j] = density_walk<false, 2>(c, s, pred_pos, nullptr, cells);
            }
            continue;
        }
        const uint32_t brick = queue[item];
        const SceneConst c = sc.scene_of_brick(brick);
        const BrickCtx bc = brick_origin(c, brick);
        halo_lookup(c, table, bc, kCapD, H, s_cellof, s_tmp);
        // Stage all runs with one load per entry; mark the entries owned by the brick.
        for (uint32_t i = threadIdx.x; i < H.staged; i += kBlock) {
            const int h = s_cellof[i];
            const uint32_t slot = (uint32_t)(H.gst[h] + ((int)i - H.oc[h].x));
            s_pos[i] = pred_pos[slot];
            s_flag[i] = interior(h) && hfull[slot] == H.hash[h] ? 1 : 0;
        }
        __syncthreads();
        for (uint32_t chunk = 0;; chunk += kOwnCap) {
            const uint32_t total = owned_round(H, s_flag, hfull, chunk, O, s_tmp);
            if (chunk >= total) break;
            const uint32_t m = min((uint32_t)kOwnCap, total - chunk);
            for (uint32_t o = threadIdx.x; o < m; o += kBlock) {
                const uint32_t j = O.slot[o];
                const uint16_t l = O.lds[o];
                float4 s;
                if (l != 0xFFFFu) s = s_pos[l];
                else s = pred_pos[j];
                const DensityCells<true> cells{CellRuns<true>{c.table_base, c.mask, table, &H, bc.ox, bc.oy, bc.oz},
                                               sample_axis(s.x, c.minx, c.maxx, c.R),
                                               sample_axis(s.y, c.miny, c.maxy, c.R),
                                               sample_axis(s.z, c.minz, c.maxz, c.R)};
                dens[j] = density_walk<true, 2>(c, s, pred_pos, s_pos, cells);
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------------------
// Force + integrate of slot j (applyPressure...comp:328-337, updateVelocities.comp:71-119).
template <bool kHaloLds>
__device__ __forceinline__ void force_of(const SceneConst &c, uint32_t j,
                                         const float4 *__restrict__ pred_pos,
                                         const float4 *__restrict__ pos_sorted,
                                         const float4 *__restrict__ vel_sorted,
                                         const float2 *__restrict__ dens,
                                         const int2 *__restrict__ table, const Halo *H,
                                         const BrickCtx &bc, const float4 *s_a, const float4 *s_b,
                                         uint32_t *q, float4 *__restrict__ pos_out,
                                         float4 *__restrict__ vel_out,
                                         float4 *__restrict__ dbg_vel_force) {
    const float4 s = pred_pos[j];
    const float4 P = pos_sorted[j];
    const float4 V = vel_sorted[j];
    const float2 myd = dens[j];
    const ForceSelf me = force_self(c, j, s, P, V, myd);
    const ForceCells<kHaloLds> cells{CellRuns<kHaloLds>{c.table_base, c.mask, table, H, bc.ox, bc.oy, bc.oz},
                                     force_cell(c, s)};
    float3 Fp = make_float3(0.0f, 0.0f, 0.0f), Fv = make_float3(0.0f, 0.0f, 0.0f);
    force_walk<kHaloLds, 2>(c, me, pred_pos, vel_sorted, dens, s_a, s_b, cells, q, kBlock, Fp, Fv);
    float4 po, vo, vf;
    force_finish(c, P, V, myd.x, Fp, Fv, po, vo, vf);
    if (dbg_vel_force) dbg_vel_force[j] = vf;
    pos_out[j] = po;
    vel_out[j] = vo;
}

template <class Scenes>
__global__ __launch_bounds__(kBlock) void k_force_brick(
    Scenes sc, const uint32_t *__restrict__ queue, uint32_t *__restrict__ counters,
    const float4 *__restrict__ pred_pos, const float4 *__restrict__ pos_sorted,
    const float4 *__restrict__ vel_sorted, const float2 *__restrict__ dens,
    const int2 *__restrict__ table, const uint32_t *__restrict__ hfull,
    const uint32_t *__restrict__ stray, float4 *__restrict__ pos_out,
    float4 *__restrict__ vel_out, float4 *__restrict__ dbg_vel_force) {
    __shared__ Halo H;
    __shared__ Owned O;
    __shared__ float4 s_a[kCapF];
    __shared__ float4 s_b[kCapF];
    __shared__ uint8_t s_cellof[kCapF];
    __shared__ uint8_t s_flag[kCapF];
    __shared__ uint32_t s_queue[kQueue * kBlock];
    __shared__ uint32_t s_tmp[kBlock / 64];
    __shared__ uint32_t s_item;
    const uint32_t qn = counters[0], ns = counters[1];
    const uint32_t items = qn + (ns + kBlock - 1u) / kBlock;
    for (;;) {
        const uint32_t item = next_item(&counters[3], &s_item);
        if (item >= items) break;
        if (item >= qn) {
            const uint32_t idx = (item - qn) * kBlock + threadIdx.x;
            if (idx < ns) {
                const uint32_t j = stray[idx];
                force_of<false>(sc.scene(j), j, pred_pos, pos_sorted, vel_sorted, dens, table,
                                nullptr, BrickCtx{0, 0, 0}, nullptr, nullptr, s_queue + threadIdx.x,
                                pos_out, vel_out, dbg_vel_force);
            }
            continue;
        }
        const uint32_t brick = queue[item];
        const SceneConst c = sc.scene_of_brick(brick);
        const BrickCtx bc = brick_origin(c, brick);
        halo_lookup(c, table, bc, kCapF, H, s_cellof, s_tmp);
        for (uint32_t i = threadIdx.x; i < H.staged; i += kBlock) {
            const int h = s_cellof[i];
            const uint32_t slot = (uint32_t)(H.gst[h] + ((int)i - H.oc[h].x));
            const float4 p = pred_pos[slot];
            const float4 v = vel_sorted[slot];
            const float2 d = dens[slot];
            s_a[i] = make_float4(p.x, p.y, p.z, d.y);
            s_b[i] = make_float4(v.x, v.y, v.z, d.x);
            s_flag[i] = interior(h) && hfull[slot] == H.hash[h] ? 1 : 0;
        }
        __syncthreads();
        for (uint32_t chunk = 0;; chunk += kOwnCap) {
            const uint32_t total = owned_round(H, s_flag, hfull, chunk, O, s_tmp);
            if (chunk >= total) break;
            const uint32_t m = min((uint32_t)kOwnCap, total - chunk);
            for (uint32_t o = threadIdx.x; o < m; o += kBlock)
                force_of<true>(c, O.slot[o], pred_pos, pos_sorted, vel_sorted, dens, table, &H, bc,
                               s_a, s_b, s_queue + threadIdx.x, pos_out, vel_out, dbg_vel_force);
            __syncthreads();
        }
    }
}

// Workgroups of `kernel` that fit on the device at once (queried outside graph capture by
// brick_prepare and cached).
template <class K>
uint32_t resident_blocks(K kernel) {
    int dev = 0, cus = 0, per_cu = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    return (uint32_t)(cus * per_cu);
}

template <class Scenes>
struct BrickGrids {
    static uint32_t density, force;
};
template <class Scenes> uint32_t BrickGrids<Scenes>::density = 0;
template <class Scenes> uint32_t BrickGrids<Scenes>::force = 0;

inline uint32_t clamp_grid(uint32_t resident, uint64_t max_items) {
    // At least 8 workgroups per CU are launched; the ones that cannot be resident start
    // as others retire and find the queue drained or not, which costs nothing either way.
    uint64_t g = resident < 2048u ? 2048u : resident;
    if (g > max_items) g = max_items;
    return g ? (uint32_t)g : 1u;
}

}  // namespace

void brick_prepare() {
    if (!BrickGrids<OneScene>::density) {
        BrickGrids<OneScene>::density = resident_blocks(k_density_brick<OneScene>);
        BrickGrids<OneScene>::force = resident_blocks(k_force_brick<OneScene>);
        BrickGrids<ManyScenes>::density = resident_blocks(k_density_brick<ManyScenes>);
        BrickGrids<ManyScenes>::force = resident_blocks(k_force_brick<ManyScenes>);
    }
}

hipError_t launch_queue(hipStream_t stream, const ChainBuffers &b, uint32_t brick_total) {
    const uint32_t blocks = (brick_total + kBlock - 1u) / kBlock;
    hipLaunchKernelGGL(k_queue, dim3(blocks ? blocks : 1u), dim3(kBlock), 0, stream,
                       b.brick_flag, brick_total, b.queue, b.counters);
    return hipGetLastError();
}

template <class Scenes>
hipError_t launch_density_brick(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                                uint32_t n, uint32_t brick_total) {
    const uint32_t grid = clamp_grid(BrickGrids<Scenes>::density,
                                     (uint64_t)brick_total + (n + kBlock - 1u) / kBlock);
    hipLaunchKernelGGL(k_density_brick<Scenes>, dim3(grid), dim3(kBlock), 0, stream, sc, b.queue,
                       b.counters, b.pred_pos, b.table, b.hfull, b.stray, b.dens);
    return hipGetLastError();
}

template <class Scenes>
hipError_t launch_force_brick(hipStream_t stream, const Scenes &sc, const ChainBuffers &b,
                              uint32_t n, uint32_t brick_total) {
    const uint32_t grid = clamp_grid(BrickGrids<Scenes>::force,
                                     (uint64_t)brick_total + (n + kBlock - 1u) / kBlock);
    hipLaunchKernelGGL(k_force_brick<Scenes>, dim3(grid), dim3(kBlock), 0, stream, sc, b.queue,
                       b.counters, b.pred_pos, b.pos_sorted, b.vel_sorted, b.dens, b.table,
                       b.hfull, b.stray, b.pos, b.vel, b.dbg_vel_force);
    return hipGetLastError();
}

template hipError_t launch_density_brick<OneScene>(hipStream_t, const OneScene &, const ChainBuffers &, uint32_t, uint32_t);
template hipError_t launch_density_brick<ManyScenes>(hipStream_t, const ManyScenes &, const ChainBuffers &, uint32_t, uint32_t);
template hipError_t launch_force_brick<OneScene>(hipStream_t, const OneScene &, const ChainBuffers &, uint32_t, uint32_t);
template hipError_t launch_force_brick<ManyScenes>(hipStream_t, const ManyScenes &, const ChainBuffers &, uint32_t, uint32_t);

}  // namespace sph
