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
constexpr int kH1 = kBrick + 2;
constexpr int kHalo = kH1 * kH1 * kH1;
constexpr int kCells = kBrick * kBrick * kBrick;
constexpr int kCapD = 2048;    // staged particles, density pass: 32 KiB of float4
constexpr int kCapF = 1280;    // staged particles, force pass: 40 B each, 50 KiB
constexpr int kOwnCap = 1024;  // particles evaluated per round of a brick
static_assert(kHalo <= kBlock && kCells <= kBlock, "one thread per halo cell");

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

struct Halo {
    int off[kHalo];  // LDS offset of the staged run; -1 = run read from global memory
    int cnt[kHalo];  // run length (0 = empty cell)
    int gst[kHalo];  // first slot of the run
};

struct BrickCtx {
    int ox, oy, oz;  // cell coordinate of halo cell (0,0,0)
};

__device__ __forceinline__ BrickCtx brick_origin(const SceneConst &c, uint32_t brick) {
    const uint32_t lb = brick - c.brick_base;
    const uint32_t bx = lb % c.nbx, by = (lb / c.nbx) % c.nby, bz = lb / (c.nbx * c.nby);
    return BrickCtx{(int)bx * kBrick - 1, (int)by * kBrick - 1, (int)bz * kBrick - 1};
}

// Looks up the halo runs and assigns LDS space in halo order until `cap` is used up.
__device__ __forceinline__ void halo_lookup(const SceneConst &c, const int2 *__restrict__ table,
                                            const BrickCtx &bc, int cap, Halo &H, uint32_t *s_tmp) {
    const int tid = threadIdx.x;
    uint32_t cnt = 0;
    int gst = -1;
    if (tid < kHalo) {
        const int hx = tid % kH1, hy = (tid / kH1) % kH1, hz = tid / (kH1 * kH1);
        const int2 se = table[key_of_cell(c, bc.ox + hx, bc.oy + hy, bc.oz + hz)];
        if (se.x != -1) {
            cnt = (uint32_t)(se.y - se.x);
            gst = se.x;
        }
    }
    uint32_t total;
    const uint32_t off = scan256(cnt, s_tmp, total);
    if (tid < kHalo) {
        H.cnt[tid] = (int)cnt;
        H.gst[tid] = gst;
        H.off[tid] = off + cnt <= (uint32_t)cap ? (int)off : -1;
    }
    __syncthreads();
}

// Number of slots of brick cell `t` (thread t < kCells) whose current cell is that cell.
struct OwnedCell {
    int gst, cnt;
    uint32_t hash;
};

__device__ __forceinline__ OwnedCell owned_cell(const BrickCtx &bc, const Halo &H) {
    const int t = threadIdx.x;
    const int lx = t % kBrick + 1, ly = (t / kBrick) % kBrick + 1, lz = t / (kBrick * kBrick) + 1;
    const int h = lx + kH1 * (ly + kH1 * lz);
    OwnedCell oc;
    oc.gst = H.gst[h];
    oc.cnt = H.cnt[h];
    oc.hash = cell_hash((uint32_t)(bc.ox + lx), (uint32_t)(bc.oy + ly), (uint32_t)(bc.oz + lz));
    return oc;
}

// Runs `eval(slot)` for every particle owned by the brick, kOwnCap at a time.
template <class Eval>
__device__ __forceinline__ void for_owned(const BrickCtx &bc, const Halo &H,
                                          const uint32_t *__restrict__ hfull, uint32_t *s_own,
                                          uint32_t *s_tmp, Eval eval) {
    const int tid = threadIdx.x;
    OwnedCell oc{0, 0, 0u};
    uint32_t own = 0;
    if (tid < kCells) {
        oc = owned_cell(bc, H);
        for (int k = 0; k < oc.cnt; ++k) own += hfull[oc.gst + k] == oc.hash ? 1u : 0u;
    }
    uint32_t total;
    const uint32_t base = scan256(own, s_tmp, total);
    for (uint32_t chunk = 0; chunk < total; chunk += kOwnCap) {
        if (tid < kCells && own) {
            uint32_t idx = base;
            for (int k = 0; k < oc.cnt; ++k) {
                if (hfull[oc.gst + k] != oc.hash) continue;
                if (idx >= chunk && idx < chunk + kOwnCap) s_own[idx - chunk] = (uint32_t)(oc.gst + k);
                ++idx;
            }
        }
        __syncthreads();
        const uint32_t m = min((uint32_t)kOwnCap, total - chunk);
        for (uint32_t o = tid; o < m; o += kBlock) eval(s_own[o]);
        __syncthreads();
    }
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

// ---------------------------------------------------------------------------------------
template <class Scenes>
__global__ __launch_bounds__(kBlock) void k_density_brick(
    Scenes sc, const uint32_t *__restrict__ queue, uint32_t *__restrict__ counters,
    const float4 *__restrict__ pred_pos, const int2 *__restrict__ table,
    const uint32_t *__restrict__ hfull, const uint32_t *__restrict__ stray,
    float2 *__restrict__ dens) {
    __shared__ Halo H;
    __shared__ float4 s_pos[kCapD];
    __shared__ uint32_t s_own[kOwnCap];
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
                const AxisCells ax = sample_axis(s.x, c.minx, c.maxx, c.R);
                const AxisCells ay = sample_axis(s.y, c.miny, c.maxy, c.R);
                const AxisCells az = sample_axis(s.z, c.minz, c.maxz, c.R);
                float M = 0.0f, Mn = 0.0f;
                for (int dz = 0; dz < 3; ++dz)
                    for (int dy = 0; dy < 3; ++dy)
                        for (int dx = 0; dx < 3; ++dx) {
                            if (!(ax.in(dx) && ay.in(dy) && az.in(dz))) continue;
                            const int2 se = table[key_of_cell(c, ax.c(dx), ay.c(dy), az.c(dz))];
                            if (se.x == -1) continue;
                            density_run(c, pred_pos + se.x, se.y - se.x, s, M, Mn);
                        }
                dens[j] = density_finish(c, M, Mn);
            }
            continue;
        }
        const SceneConst c = sc.scene_of_brick(queue[item]);
        const BrickCtx bc = brick_origin(c, queue[item]);
        halo_lookup(c, table, bc, kCapD, H, s_tmp);
        if (threadIdx.x < kHalo && H.off[threadIdx.x] >= 0) {
            const int o = H.off[threadIdx.x], g = H.gst[threadIdx.x], n = H.cnt[threadIdx.x];
            for (int k = 0; k < n; ++k) s_pos[o + k] = pred_pos[g + k];
        }
        __syncthreads();
        for_owned(bc, H, hfull, s_own, s_tmp, [&](uint32_t j) {
            const float4 s = pred_pos[j];
            const AxisCells ax = sample_axis(s.x, c.minx, c.maxx, c.R);
            const AxisCells ay = sample_axis(s.y, c.miny, c.maxy, c.R);
            const AxisCells az = sample_axis(s.z, c.minz, c.maxz, c.R);
            float M = 0.0f, Mn = 0.0f;
            for (int dz = 0; dz < 3; ++dz)
                for (int dy = 0; dy < 3; ++dy)
                    for (int dx = 0; dx < 3; ++dx) {
                        if (!(ax.in(dx) && ay.in(dy) && az.in(dz))) continue;
                        const int lx = ax.c(dx) - bc.ox, ly = ay.c(dy) - bc.oy, lz = az.c(dz) - bc.oz;
                        if ((unsigned)lx < (unsigned)kH1 && (unsigned)ly < (unsigned)kH1 &&
                            (unsigned)lz < (unsigned)kH1) {
                            const int h = lx + kH1 * (ly + kH1 * lz);
                            const int n = H.cnt[h];
                            if (n == 0) continue;
                            const int o = H.off[h];
                            if (o >= 0) density_run(c, s_pos + o, n, s, M, Mn);
                            else density_run(c, pred_pos + H.gst[h], n, s, M, Mn);
                        } else {
                            const int2 se = table[key_of_cell(c, ax.c(dx), ay.c(dy), az.c(dz))];
                            if (se.x == -1) continue;
                            density_run(c, pred_pos + se.x, se.y - se.x, s, M, Mn);
                        }
                    }
            dens[j] = density_finish(c, M, Mn);
        });
    }
}

// ---------------------------------------------------------------------------------------
template <class Scenes>
__device__ __forceinline__ void force_one_global(const SceneConst &c, uint32_t j,
                                                 const float4 *__restrict__ pred_pos,
                                                 const float4 *__restrict__ pos_sorted,
                                                 const float4 *__restrict__ vel_sorted,
                                                 const float2 *__restrict__ dens,
                                                 const int2 *__restrict__ table,
                                                 float4 *__restrict__ pos_out,
                                                 float4 *__restrict__ vel_out,
                                                 float4 *__restrict__ dbg_vel_force) {
    const float4 s = pred_pos[j];
    const float4 P = pos_sorted[j];
    const float4 V = vel_sorted[j];
    const float2 myd = dens[j];
    const ForceSelf me = force_self(c, j, s, P, V, myd);
    const int3 c0 = force_cell(c, s);
    float3 Fp = make_float3(0.0f, 0.0f, 0.0f), Fv = make_float3(0.0f, 0.0f, 0.0f);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int2 se = table[key_of_cell(c, c0.x + dx, c0.y + dy, c0.z + dz)];
                if (se.x == -1) continue;
                force_run(c, pred_pos + se.x, vel_sorted + se.x, dens + se.x, se.y - se.x,
                          (uint32_t)se.x, me, Fp, Fv);
            }
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
    __shared__ float4 s_pos[kCapF];
    __shared__ float4 s_vel[kCapF];
    __shared__ float2 s_dd[kCapF];
    __shared__ uint32_t s_own[kOwnCap];
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
                force_one_global<Scenes>(sc.scene(j), j, pred_pos, pos_sorted, vel_sorted, dens,
                                         table, pos_out, vel_out, dbg_vel_force);
            }
            continue;
        }
        const SceneConst c = sc.scene_of_brick(queue[item]);
        const BrickCtx bc = brick_origin(c, queue[item]);
        halo_lookup(c, table, bc, kCapF, H, s_tmp);
        if (threadIdx.x < kHalo && H.off[threadIdx.x] >= 0) {
            const int o = H.off[threadIdx.x], g = H.gst[threadIdx.x], n = H.cnt[threadIdx.x];
            for (int k = 0; k < n; ++k) {
                s_pos[o + k] = pred_pos[g + k];
                s_vel[o + k] = vel_sorted[g + k];
                s_dd[o + k] = dens[g + k];
            }
        }
        __syncthreads();
        for_owned(bc, H, hfull, s_own, s_tmp, [&](uint32_t j) {
            const float4 s = pred_pos[j];
            const float4 P = pos_sorted[j];
            const float4 V = vel_sorted[j];
            const float2 myd = dens[j];
            const ForceSelf me = force_self(c, j, s, P, V, myd);
            const int3 c0 = force_cell(c, s);
            float3 Fp = make_float3(0.0f, 0.0f, 0.0f), Fv = make_float3(0.0f, 0.0f, 0.0f);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int cx = c0.x + dx, cy = c0.y + dy, cz = c0.z + dz;
                        const int lx = cx - bc.ox, ly = cy - bc.oy, lz = cz - bc.oz;
                        if ((unsigned)lx < (unsigned)kH1 && (unsigned)ly < (unsigned)kH1 &&
                            (unsigned)lz < (unsigned)kH1) {
                            const int h = lx + kH1 * (ly + kH1 * lz);
                            const int n = H.cnt[h];
                            if (n == 0) continue;
                            const int o = H.off[h];
                            const uint32_t g = (uint32_t)H.gst[h];
                            if (o >= 0) force_run(c, s_pos + o, s_vel + o, s_dd + o, n, g, me, Fp, Fv);
                            else force_run(c, pred_pos + g, vel_sorted + g, dens + g, n, g, me, Fp, Fv);
                        } else {
                            const int2 se = table[key_of_cell(c, cx, cy, cz)];
                            if (se.x == -1) continue;
                            force_run(c, pred_pos + se.x, vel_sorted + se.x, dens + se.x,
                                      se.y - se.x, (uint32_t)se.x, me, Fp, Fv);
                        }
                    }
            float4 po, vo, vf;
            force_finish(c, P, V, myd.x, Fp, Fv, po, vo, vf);
            if (dbg_vel_force) dbg_vel_force[j] = vf;
            pos_out[j] = po;
            vel_out[j] = vo;
        });
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
    uint64_t g = resident ? resident : 256u;
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
