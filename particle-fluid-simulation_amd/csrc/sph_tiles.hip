// Tiled density and force passes: calculateDensities.comp:96-155 and
// applyPressureAndViscosityForce.comp:275-337 + updateVelocities.comp:71-119 over cell tiles
// staged in LDS.
//
// One workgroup per tile (sph_tile_owner.hpp).  It looks up the cell-ordered run of every
// cell of the tile's region (the tile grown by one cell per side) in the halo grid, stages
// the runs in LDS in region order (x fastest), and selects the particles the tile owns from
// the staged copies.  For an owned particle the three cells (x - 1 .. x + 1) of each of its
// nine (dz, dy) rows are consecutive region cells, so each row of the 27-cell walk is one
// contiguous LDS range, visited in the reference's order.  A region larger than the LDS
// capacity streams through it in groups of rows (see plan_groups).
//
// Density: a row is first scanned with the distance test only, and the candidates that
// pass (d2 <= R2_pre) are queued per lane in LDS; the queue is drained in the order it was
// filled (after a row once some lane holds kTFlush entries, when a lane's queue is full,
// and at the end) with the density terms and the force-list entries.  Every particle's sums
// thus take the reference's candidates in the reference's order, bit for bit, while the
// in-range arithmetic runs with most lanes of the wave active.  The force-list entries are
// either cell-ordered slots, read by the per-lane force pass, or 16-bit indices into the
// staged region, read by the tiled force pass, which stages the same region (position +
// near density, velocity + density per entry) and reads its neighbours from LDS.
//
// The particles no tile owns (the placement's strays) take one lane each in the first
// workgroups of each tiled launch, walking their cells through global memory beside the
// tiles.  A tile that cannot be staged (see plan_groups) hands its particles over to the top
// of the stray list, which a stray kernel after each tiled pass walks the same way.
#include <algorithm>

#include "sph_device.hpp"
#include "sph_kernels.hpp"
#include "sph_pairs.hpp"
#include "sph_tile_owner.hpp"

namespace sph {

namespace {

constexpr int kTQ = 24;     // density candidate queue entries per lane
constexpr int kTFlush = 16; // drain once some lane holds this many

// Lanes of a tiled workgroup that walk particles through global memory (strays, tiles that
// cannot be staged): their list entries are staged in the tile's LDS copy of positions.
template <int kB>
constexpr int kGlobalLanes = kB < 256 ? kB : 256;

// Tile of the r-th tile workgroup (r counted from a multiple of 8): tiles dealt to the 8
// XCDs in chunks of kTileChunk consecutive tiles, XCD x taking chunks x, x + 8, ...  On the
// 1M fill, one contiguous range of tiles per XCD put the dense floor and wall tiles on a
// few XCDs (241 us); launch order balanced them (193 us) but read 87 MiB per launch from
// HBM; chunks of 4 keep the time with 70 MiB (16: 195 us, 66 MiB; 64: 218 us, 47 MiB).
constexpr uint32_t kTileChunk = 4;
__device__ __forceinline__ uint32_t tile_of_block(uint32_t r, uint32_t nt) {
    constexpr uint32_t kChunk = kTileChunk, kSpan = 8u * kChunk;
    if (r >= nt / kSpan * kSpan) return r;
    const uint32_t i = r / 8u;
    return ((i / kChunk) * 8u + r % 8u) * kChunk + i % kChunk;
}

__device__ __forceinline__ GridRuns tile_grid_runs(const SceneConst &c, const ChainBuffers &b) {
    return GridRuns{b.grid, c.grid_base, (uint32_t)c.gx + 2u, (uint32_t)c.gy + 2u,
                    c.grid_cells ? (uint32_t)c.gz + 2u : 0u, c.table_base, c.mask, b.crun};
}

// Looks up the run of every region cell: s_beg[rc] = its first cell-ordered slot,
// s_off[rc] = exclusive prefix of the run lengths in region order, s_off[cells] = total.
// Ends with a barrier; returns the total.
template <int kB>
__device__ __forceinline__ uint32_t region_runs(const SceneConst &c, const ChainBuffers &b, const TileRegion &rg,
                                                int *s_off, int *s_beg, uint32_t *s_wave) {
    constexpr uint32_t kPer = (kTileRegionMax + kB - 1u) / kB;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const GridRuns runs = tile_grid_runs(c, b);
    int cnt[kPer];
    uint32_t sum = 0;
    // region coordinates of the lane's first cell, then stepped x-fastest
    uint32_t cxr = (tid * kPer) % rg.rx, yz = (tid * kPer) / rg.rx;
    uint32_t cyr = yz % rg.ry, czr = yz / rg.ry;
    int2 se[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        se[k] = make_int2(-1, -1);
        if (tid * kPer + k < rg.cells)
            se[k] = runs.raw(rg.ox + (int32_t)cxr, rg.oy + (int32_t)cyr, rg.oz + (int32_t)czr);
        if (++cxr == rg.rx) {
            cxr = 0;
            if (++cyr == rg.ry) {
                cyr = 0;
                ++czr;
            }
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t rc = tid * kPer + k;
        int len = 0;
        if (rc < rg.cells && se[k].x >= 0) {
            len = se[k].y - se[k].x;
            s_beg[rc] = se[k].x;
        }
        cnt[k] = len;
        sum += (uint32_t)len;
    }
    uint32_t x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63u) s_wave[w] = x;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
#pragma unroll
    for (uint32_t i = 0; i < (uint32_t)kB / 64u; ++i) {
        if (i < w) wbase += s_wave[i];
        total += s_wave[i];
    }
    uint32_t run = wbase + x - sum;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t rc = tid * kPer + k;
        if (rc < rg.cells) s_off[rc] = (int)run;
        run += (uint32_t)cnt[k];
    }
    if (tid == 0) s_off[rg.cells] = (int)total;
    __syncthreads();
    return total;
}

// Squared distance for the density pass's candidate filter: fused multiply-adds, within 6
// ulp of dist2 (each of the two evaluations rounds three times, every term non-negative).
__device__ __forceinline__ float dist2_filter(const float4 &p, const float4 &s) {
    const float ex = p.x - s.x, ey = p.y - s.y, ez = p.z - s.z;
    return fmaf(ez, ez, fmaf(ey, ey, ex * ex));
}

// Records the particle as owned when this staged copy is its home copy: one LDS atomic per
// wave, so a wave's owned particles keep their region order.
template <int kOwnMax>
__device__ __forceinline__ void select_own(bool own, uint32_t e, uint32_t g, uint32_t *s_count,
                                           uint16_t *s_own_e, uint32_t *s_own_g) {
    const uint64_t m = __ballot(own);
    if (!m) return;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(s_count, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader, 64);
    const uint32_t i = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (own && i < (uint32_t)kOwnMax) {
        s_own_e[i] = (uint16_t)e;
        if (s_own_g) s_own_g[i] = g;
    }
}

// Density of slot j through global memory (strays, tiles too large for LDS); no force list.
template <bool kF>
__device__ __forceinline__ void density_global(const SceneConst &c, const ChainBuffers &b, uint32_t j) {
    const float4 s = b.pred_pos[j];
    const DensityCellsOf<GridRuns> cells{tile_grid_runs(c, b), sample_axis(s.x, c.minx, c.maxx, c.R),
                                         sample_axis(s.y, c.miny, c.maxy, c.R),
                                         sample_axis(s.z, c.minz, c.maxz, c.R)};
    const float2 d = density_walk_rows<2, kF>(c, s, b.pred_pos, cells);
    b.dens[j] = d;
    reinterpret_cast<float *>(b.vel_sorted + j)[3] = d.x;
    reinterpret_cast<float *>(b.pred_pos + j)[3] = d.y;
    b.ncount[j] = kListInvalid;
}

// The longest force list above half the allocated capacity, for the host's list sizing
// (as the per-lane density pass records it).
__device__ __forceinline__ void record_longest(const ChainBuffers &b, uint32_t found) {
    const uint64_t longer = __ballot(found > (b.list_cap >> 1));
    if (!longer) return;
    uint32_t longest = 0;
    for (uint64_t m = longer; m; m &= m - 1ull)
        longest = max(longest, (uint32_t)__builtin_amdgcn_readlane((int)found, __ffsll((long long)m) - 1));
    if ((threadIdx.x & 63u) == (uint32_t)(__ffsll((long long)longer) - 1) &&
        __hip_atomic_load(&b.counters[4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < longest)
        atomicMax(&b.counters[4], longest);
}

// Density of slot j through global memory, recording its force list of cell-ordered slots
// as the per-lane density pass does (lane_list: kListLds staged entries at `stride`), for
// the strays of launches whose force pass reads slot lists.
template <bool kF>
__device__ __forceinline__ void density_listed(const SceneConst &c, const ChainBuffers &b, uint32_t n, uint32_t j,
                                               uint32_t *lane_list, int stride) {
    const float4 s = b.pred_pos[j];
    uint32_t count, found;
    const float2 d = density_with_list<4, kF>(c, j, s, b.pred_pos, tile_grid_runs(c, b), lane_list, stride, b.nlist,
                                              n, b.list_cap, count, found);
    record_longest(b, found);
    b.dens[j] = d;
    reinterpret_cast<float *>(b.vel_sorted + j)[3] = d.x;
    reinterpret_cast<float *>(b.pred_pos + j)[3] = d.y;
    b.ncount[j] = (uint8_t)count;
    const uint32_t m = count == kListInvalid ? 0u : min(count, (uint32_t)kListLds);
    for (uint32_t k = 0; __ballot(k < m); ++k)
        if (k < m) b.nlist[(size_t)k * n + j] = lane_list[k * stride];
}

// Force terms of slot j whose neighbours are read from global memory: the 27-cell walk
// queueing in-range candidates in LDS (q, stride `stride`) as the per-lane kernels do, for
// strays, tiles too large for LDS and lists that did not fit.
struct ForceTerms {
    float3 Fp, Fv;
};
template <bool kF>
__device__ __forceinline__ ForceTerms force_global(const SceneConst &c, const ChainBuffers &b, const ForceSelf &me,
                                                   uint32_t *q, int stride) {
    ForceTerms t{make_float3(0.0f, 0.0f, 0.0f), make_float3(0.0f, 0.0f, 0.0f)};
    const ForceCellsOf<GridRuns> cells{tile_grid_runs(c, b), force_cell(c, me.s)};
    force_walk_rows<2, kF>(c, me, b.pred_pos, b.vel_sorted, b.dens, cells, q, stride, t.Fp, t.Fv);
    return t;
}

// Integrate slot j after its force terms and store the new state at its hash slot.
__device__ __forceinline__ void force_store(const SceneConst &c, const ChainBuffers &b, uint32_t j, const float4 &P,
                                            const float4 &V, float density, const float3 &Fp, const float3 &Fv) {
    float4 po, vo, vf;
    force_finish(c, P, V, density, Fp, Fv, po, vo, vf);
    const uint32_t o = b.hslot[j];
    if (b.dbg_vel_force) {
        vf.w = b.dbg_vel_force[j].w + 1.0f;
        b.dbg_vel_force[j] = vf;
    }
    b.pos[o] = po;
    b.vel[o] = vo;
}

// Row groups of a tile.  The region is staged in LDS one group of consecutive region rows
// (a row: the rx cells of one (y, z), x fastest; rows numbered y fastest) at a time, as
// many rows as fit the scene's tile_emax entries.  A particle's nine walk rows have
// increasing region-row indices in the walk's order ((cz - 1 + dz) ry + (cy - 1 + dy), with
// ry >= 3), so walking each group's rows in turn, the sums carried from group to group,
// keeps the reference's order while every owned particle walks every group.  A row that
// alone does not fit, more owned particles than kOwnMax, or more than one group where the
// force lists index the staging hands the tile's particles over to the stray list.
constexpr int kGroupMax = 64;  // region rows (th + 2) (td + 2) a tile may have

struct TileLds {
    int s_off[kTileRegionMax + 1];  // exclusive prefix of the region cells' run lengths
    int s_beg[kTileRegionMax];      // first cell-ordered slot of each region cell's run
    uint8_t s_g0[kGroupMax + 1];    // row groups: group k holds rows [s_g0[k], s_g0[k + 1])
    uint32_t s_wave[16];
    uint32_t s_ng, s_count;
};

// Row groups of a region of `total` entries (s_off filled): their count, 0 when the tile is
// handed over (a row alone exceeds emax, or `single` and the region does not fit at once).
// Clears the owned count; ends with a barrier.
template <int kB>
__device__ __forceinline__ uint32_t plan_groups(const TileRegion &rg, int total, int emax, bool single, TileLds &L) {
    if (threadIdx.x == 0) {
        L.s_count = 0u;
        const uint32_t rows = rg.ry * rg.rz;
        uint32_t ng = 0;
        L.s_g0[0] = 0;
        if (total <= emax) {
            ng = 1;
        } else if (!single) {
            int base = 0;  // first entry of the open group
            for (uint32_t r = 0; r < rows; ++r) {
                const int lo = L.s_off[r * rg.rx], hi = L.s_off[(r + 1u) * rg.rx];
                if (hi - lo > emax) {
                    ng = 0;
                    break;
                }
                if (hi - base > emax) {  // row r opens the next group
                    L.s_g0[++ng] = (uint8_t)r;
                    base = lo;
                }
                if (r + 1u == rows) ++ng;
            }
        }
        if (ng) L.s_g0[ng] = (uint8_t)rows;
        L.s_ng = ng;
    }
    __syncthreads();
    return L.s_ng;
}

// Region cell of region entry e among cells [c0, c1) (s_off[c0] <= e < s_off[c1]).
__device__ __forceinline__ uint32_t region_cell_in(const int *s_off, uint32_t c0, uint32_t c1, int e) {
    while (c1 - c0 > 1u) {
        const uint32_t mid = (c0 + c1) >> 1;
        if (s_off[mid] <= e) c0 = mid; else c1 = mid;
    }
    return c0;
}

// Stages region entries [e0, e1) (those of cells [c0, c1)) through `stage(k, g)` at staged
// index k = e - e0 (g: the entry's cell-ordered slot).  With kSelect, also records the home
// copies of the tile's owned particles: s_own_e their region entry, s_own_g (if given) their
// slot.  Ends with a barrier.
template <int kB, int kOwnMax, bool kSelect, class Stage>
__device__ __forceinline__ void stage_group(const ChainBuffers &b, uint32_t t, uint32_t c0, uint32_t c1, int e0,
                                            int e1, TileLds &L, uint16_t *s_own_e, uint32_t *s_own_g,
                                            Stage stage) {
    constexpr int kPer = 5;  // a lane's loads are issued before any is used
    const int tid = (int)threadIdx.x;
    const int cnt = e1 - e0;
    for (int k0 = 0; k0 < cnt; k0 += kPer * kB) {
        uint32_t rc[kPer], g[kPer];
        uint2 ow[kPer];
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = k0 + tid + u * kB;
            rc[u] = 0u;
            g[u] = 0u;
            if (k < cnt) {
                rc[u] = region_cell_in(L.s_off, c0, c1, e0 + k);
                g[u] = (uint32_t)(L.s_beg[rc[u]] + (e0 + k - L.s_off[rc[u]]));
            }
        }
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = k0 + tid + u * kB;
            if (k < cnt) {
                stage(k, g[u]);
                if (kSelect) ow[u] = b.towner[g[u]];
            }
        }
        if (kSelect) {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int k = k0 + tid + u * kB;
                if (k - tid >= cnt) break;  // the same for every lane of the block
                const bool own = k < cnt && ow[u].x == t && ow[u].y == rc[u];
                select_own<kOwnMax>(own, (uint32_t)(e0 + k), g[u], &L.s_count, s_own_e, s_own_g);
            }
        }
    }
    __syncthreads();
}

// Hands the tile's owned particles over to the stray list, from its top (the placement's
// strays fill it from the bottom).
template <int kB>
__device__ __forceinline__ void hand_over(const ChainBuffers &b, uint32_t n, uint32_t t, const TileRegion &rg,
                                          int total, const TileLds &L) {
    const uint32_t tid = threadIdx.x;
    for (int e0 = 0; e0 < total; e0 += kB) {
        const int e = e0 + (int)tid;
        uint32_t j = kNoTile;
        if (e < total) {
            const uint32_t rc = region_cell_in(L.s_off, 0u, rg.cells, e);
            const uint32_t g = (uint32_t)(L.s_beg[rc] + (e - L.s_off[rc]));
            const uint2 ow = b.towner[g];
            if (ow.x == t && ow.y == rc) j = g;
        }
        const bool mine = j != kNoTile;
        const uint64_t m = __ballot(mine);
        if (!m) continue;
        const uint32_t lane = tid & 63u;
        const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&b.counters[7], (uint32_t)__popcll(m));
        base = __shfl(base, (int)leader, 64);
        if (mine) b.stray[n - 1u - (base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)))] = j;
    }
}

// A tile that cannot be staged, with slot lists: its workgroup walks its owned particles
// itself, one lane each, through global memory, recording their lists as the per-lane
// density pass does.  The region's entries are scanned in rounds of kRound; each round's
// owned slots are collected in `s_g` and walked with their list entries staged in
// `lane_list` (kListLds per lane).  The particles count among the strays (counters[7]).
template <int kB, int kRound, bool kF, class Scenes>
__device__ __forceinline__ void walk_tile_global(const Scenes &sc, const ChainBuffers &b, uint32_t n, uint32_t t,
                                                 const TileRegion &rg, int total, TileLds &L, uint32_t *s_g,
                                                 uint32_t *lane_list) {
    const uint32_t tid = threadIdx.x;
    for (int base = 0; base < total; base += kRound) {
        if (tid == 0) L.s_count = 0u;
        __syncthreads();
        const int end = min(total, base + kRound);
        for (int e0 = base; e0 < end; e0 += kB) {
            const int e = e0 + (int)tid;
            bool own = false;
            uint32_t g = 0;
            if (e < end) {
                const uint32_t rc = region_cell_in(L.s_off, 0u, rg.cells, e);
                g = (uint32_t)(L.s_beg[rc] + (e - L.s_off[rc]));
                const uint2 ow = b.towner[g];
                own = ow.x == t && ow.y == rc;
            }
            const uint64_t m = __ballot(own);
            if (!m) continue;
            const uint32_t lane = tid & 63u;
            const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1u;
            uint32_t at = 0;
            if (lane == leader) at = atomicAdd(&L.s_count, (uint32_t)__popcll(m));
            at = __shfl(at, (int)leader, 64);
            if (own) s_g[at + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = g;
        }
        __syncthreads();
        const uint32_t cnt = L.s_count;
        if (tid == 0 && cnt) atomicAdd(&b.counters[7], cnt);
        constexpr int kGL = kGlobalLanes<kB>;
        if (tid < (uint32_t)kGL)
            for (uint32_t i = tid; i < cnt; i += kGL) {
                const uint32_t j = s_g[i];
                density_listed<kF>(sc.scene(j), b, n, j, lane_list + tid, kGL);
            }
        __syncthreads();  // s_g and lane_list are reused by the next round
    }
}

// The density walk of one owned particle over the rows of group [r0, r1), staged from
// region entry e0: S = (predicted position xyz, slot); M, Mn, nl carry its sums and list
// length from group to group.
//
// Rows 0..8 of the walk, then a final pass (row 9, empty) that only drains; rows outside the
// group are empty.  A row's candidates are filtered with a fused-multiply-add distance
// (within 6 ulp of the reference's d2, against a bound 2^-20 above its cut-off, so no
// candidate the reference keeps is dropped) and queued without branches (the entry is
// written, the count advances when it passes), in chunks that fit the queue.  The drain
// recomputes d2 exactly and applies the reference's tests.  A wall particle drains at the
// end of each row, while that row's density cells are known.
template <int kB, bool kF, bool kS>
__device__ __forceinline__ void walk_group(const SceneConst &c, const ChainBuffers &b, uint32_t n,
                                           const TileRegion &rg, const TileLds &L, const float4 *s_pos, uint16_t *q,
                                           const float4 &S, uint32_t r0, uint32_t r1, int e0, float &M, float &Mn,
                                           uint32_t &nl) {
    const uint32_t cap = b.list_cap;
    uint16_t *__restrict__ nl16 = b.nlist16();
    const float thr = c.R2_pre * (1.0f + 0x1p-20f);
    const uint32_t j = __float_as_uint(S.w);
    const float4 s = make_float4(S.x, S.y, S.z, 1.0f);
    const AxisCells ax = sample_axis(s.x, c.minx, c.maxx, c.R);
    const AxisCells ay = sample_axis(s.y, c.miny, c.maxy, c.R);
    const AxisCells az = sample_axis(s.z, c.minz, c.maxz, c.R);
    // region coordinates of the predicted cell, each in [1, tile extent]
    const uint32_t cx = (uint32_t)((int32_t)ax.mid - rg.ox);
    const uint32_t cy = (uint32_t)((int32_t)ay.mid - rg.oy);
    const uint32_t cz = (uint32_t)((int32_t)az.mid - rg.oz);
    // interior: every sample point of the 27 cells lies inside the box, so every candidate
    // counts for the density (else only those of the row's density cells)
    const bool interior = (ax.inmask & ay.inmask & az.inmask) == 7u;
    int qc = 0;
    for (uint32_t r = 0; r <= 9u; ++r) {
        int e = 0, o3 = 0, mlo = 0, mhi = 0;
        if (r < 9u) {
            const uint32_t dz = r / 3u, dy = r % 3u;
            const uint32_t row = (cz - 1u + dz) * rg.ry + (cy - 1u + dy);
            if (row >= r0 && row < r1) {
                const uint32_t rc0 = row * rg.rx + (cx - 1u);
                const int o0 = L.s_off[rc0] - e0;
                o3 = L.s_off[rc0 + 3u] - e0;
                e = o0;
                if (!interior) {
                    // the three force cells whose sample points are inside the box (the
                    // middle one always is)
                    const bool row_in = ay.in((int)dy) && az.in((int)dz);
                    mlo = row_in ? (ax.in(0) ? o0 : L.s_off[rc0 + 1u] - e0) : o3;
                    mhi = row_in ? (ax.in(2) ? o3 : L.s_off[rc0 + 2u] - e0) : o3;
                }
            }
        }
        for (;;) {
            const bool need = qc > 0 && (r == 9u || qc >= kTFlush || (!interior && e >= o3));
            if (__ballot(need)) {
                for (int k = 0; k < qc; ++k) {
                    const int ec = q[k * kB];
                    const float4 p = s_pos[ec];
                    const float d2 = dist2(p, s);
                    // without a branch: +0 outside the row's density cells or out of range
                    density_terms_masked<kF>(c, d2, interior || (ec >= mlo && ec < mhi), M, Mn);
                    const uint32_t slot = __float_as_uint(p.w);
                    if (slot != j && !(d2 > c.R2)) {
                        if (nl < cap) {
                            // slot lists, or 16-bit region entries (one group: e0 = 0)
                            if (kS) b.nlist[(size_t)nl * n + j] = slot;
                            else nl16[(size_t)nl * n + j] = (uint16_t)ec;
                        }
                        ++nl;
                    }
                }
                qc = 0;
            }
            if (!__ballot(e < o3)) break;
            const int lim = min(o3, e + (kTQ - qc));
            for (; e + 3 < lim; e += 4) {
                float f[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) f[u] = dist2_filter(s_pos[e + u], s);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    q[qc * kB] = (uint16_t)(e + u);
                    qc += f[u] > thr ? 0 : 1;
                }
            }
            for (; e < lim; ++e) {
                q[qc * kB] = (uint16_t)e;
                qc += dist2_filter(s_pos[e], s) > thr ? 0 : 1;
            }
        }
    }
}

// kS: force lists of cell-ordered slots (for the per-lane force pass), else of 16-bit
// region entries (for the tiled one).
template <class Scenes, int kB, int kE, bool kF, bool kS>
__global__ __launch_bounds__(kB) void k_density_tile(Scenes sc, const ChainBuffers b, uint32_t n, uint32_t lead,
                                                     uint32_t tile_total) {
    constexpr int kOwnMax = 2 * kB;
    constexpr int kSlots = kOwnMax / kB;  // owned particles per lane
    __shared__ float4 s_pos[kE];  // (predicted position xyz, cell-ordered slot)
    __shared__ TileLds L;
    __shared__ uint16_t s_own[kOwnMax];
    __shared__ alignas(16) uint16_t s_q[kTQ * kB];  // also walk_tile_global's u32 round buffer
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x < lead) {
        // the placement's strays, beside the tiles (their walks' latency overlaps them); with
        // slot lists they record theirs, their list entries staged in s_pos
        constexpr int kGL = kGlobalLanes<kB>;
        static_assert(!kS || kE * 4 >= kListLds * kGL, "the strays' list staging lives in s_pos");
        const uint32_t count = b.counters[1];
        if (tid < (uint32_t)kGL)
            for (uint32_t i = blockIdx.x * kGL + tid; i < count; i += lead * kGL) {
                const uint32_t j = b.stray[i];
                if (kS) density_listed<kF>(sc.scene(j), b, n, j, reinterpret_cast<uint32_t *>(s_pos) + tid, kGL);
                else density_global<kF>(sc.scene(j), b, j);
            }
        return;
    }
    const uint32_t t = tile_of_block(blockIdx.x - lead, tile_total);
    const SceneConst c = sc.scene_of_tile(t);
    if (!c.use_tiles) return;
    const TileRegion rg = tile_region(c, t - c.tile_base);
    const int emax = min(kE, (int)c.tile_emax);
    const int total = (int)region_runs<kB>(c, b, rg, L.s_off, L.s_beg, L.s_wave);
    if (total == 0) return;
    // 16-bit entries (s_own, the queue, 16-bit lists) cover the region; 16-bit lists index
    // a single staging
    const uint32_t ng = total > 65535 ? 0u : plan_groups<kB>(rg, total, emax, !kS, L);
    // a tile that cannot be staged: with slot lists its workgroup walks its particles
    // through global memory (round buffer: the queue; list staging: s_pos); else they go
    // to the stray kernel after this launch
    constexpr int kRound = kTQ * kB / 2;
    static_assert(kE * 4 >= kListLds * kGlobalLanes<kB>, "the list staging of walk_tile_global lives in s_pos");
    if (ng == 0u) {
        if (tid == 0) atomicAdd(&b.counters[6], 1u);
        if (kS)
            walk_tile_global<kB, kRound, kF>(sc, b, n, t, rg, total, L, reinterpret_cast<uint32_t *>(s_q),
                                             reinterpret_cast<uint32_t *>(s_pos));
        else
            hand_over<kB>(b, n, t, rg, total, L);
        return;
    }
    if (tid == 0 && ng > 1u) atomicAdd(&b.counters[5], 1u);
    auto stage_pos = [&](int k, uint32_t g) {
        const float4 p = b.pred_pos[g];
        s_pos[k] = make_float4(p.x, p.y, p.z, __uint_as_float(g));
    };
    // one group: staged with the owned particles' selection; else the selection alone
    if (ng == 1u)
        stage_group<kB, kOwnMax, true>(b, t, 0u, rg.cells, 0, total, L, s_own, nullptr, stage_pos);
    else
        stage_group<kB, kOwnMax, true>(b, t, 0u, rg.cells, 0, total, L, s_own, nullptr, [](int, uint32_t) {});
    const uint32_t owned = L.s_count;
    if (owned > (uint32_t)kOwnMax) {
        if (tid == 0) atomicAdd(&b.counters[6], 1u);
        __syncthreads();  // every lane has read the owned count before the rounds reset it
        if (kS)
            walk_tile_global<kB, kRound, kF>(sc, b, n, t, rg, total, L, reinterpret_cast<uint32_t *>(s_q),
                                             reinterpret_cast<uint32_t *>(s_pos));
        else
            hand_over<kB>(b, n, t, rg, total, L);
        return;
    }
    // the lane's owned particles tid + q kB: (predicted position, slot) and running sums
    float4 S[kSlots];
    float M[kSlots], Mn[kSlots];
    uint32_t nl[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        M[q] = 0.0f;
        Mn[q] = 0.0f;
        nl[q] = 0u;
        S[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const uint32_t i = tid + (uint32_t)q * kB;
        if (i < owned) {
            const int e = s_own[i];
            if (ng == 1u) {
                S[q] = s_pos[e];
            } else {
                const uint32_t rc = region_cell_in(L.s_off, 0u, rg.cells, e);
                const uint32_t g = (uint32_t)(L.s_beg[rc] + (e - L.s_off[rc]));
                const float4 p = b.pred_pos[g];
                S[q] = make_float4(p.x, p.y, p.z, __uint_as_float(g));
            }
        }
    }
    for (uint32_t gi = 0; gi < ng; ++gi) {
        const uint32_t r0 = L.s_g0[gi], r1 = L.s_g0[gi + 1u];
        const int e0 = L.s_off[r0 * rg.rx];
        if (ng > 1u) {
            if (gi > 0u) __syncthreads();  // the previous group's walks are done with s_pos
            stage_group<kB, kOwnMax, false>(b, t, r0 * rg.rx, r1 * rg.rx, e0, L.s_off[r1 * rg.rx], L, nullptr,
                                            nullptr, stage_pos);
        }
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            if (tid + (uint32_t)q * kB >= owned) break;
            walk_group<kB, kF, kS>(c, b, n, rg, L, s_pos, s_q + tid, S[q], r0, r1, e0, M[q], Mn[q], nl[q]);
        }
    }
    const uint32_t cap = b.list_cap;
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        if (tid + (uint32_t)q * kB >= owned) break;
        const uint32_t j = __float_as_uint(S[q].w);
        const float2 d = density_finish(c, M[q], Mn[q]);
        record_longest(b, nl[q]);
        b.dens[j] = d;
        reinterpret_cast<float *>(b.vel_sorted + j)[3] = d.x;
        reinterpret_cast<float *>(b.pred_pos + j)[3] = d.y;
        b.ncount[j] = nl[q] <= cap ? (uint8_t)nl[q] : kListInvalid;
    }
}

template <class Scenes, int kB, int kE, bool kF>
__global__ __launch_bounds__(kB) void k_force_tile(Scenes sc, const ChainBuffers b, uint32_t n, uint32_t lead,
                                                   uint32_t tile_total) {
    constexpr int kOwnMax = 2 * kB;
    static_assert(kE * 4 >= kQueue * kB, "the strays' walk queue lives in s_b");
    __shared__ float4 s_a[kE];  // (predicted position xyz, near density)
    __shared__ float4 s_b[kE];  // (velocity xyz, density)
    __shared__ TileLds L;
    __shared__ uint16_t s_own_e[kOwnMax];
    __shared__ uint32_t s_own_g[kOwnMax];
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x < lead) {
        // the placement's strays, beside the tiles; their walks queue candidates in s_b
        const uint32_t count = b.counters[1];
        for (uint32_t i = blockIdx.x * kB + tid; i < count; i += lead * kB) {
            const uint32_t j = b.stray[i];
            const SceneConst cj = sc.scene(j);
            const float4 P = b.pos_sorted[j];
            const float4 V = b.vel_sorted[j];
            const float2 myd = b.dens[j];
            const ForceSelf me = force_self(cj, j, b.pred_pos[j], P, V, myd);
            const ForceTerms ft = force_global<kF>(cj, b, me, reinterpret_cast<uint32_t *>(s_b) + tid, kB);
            force_store(cj, b, j, P, V, myd.x, ft.Fp, ft.Fv);
        }
        return;
    }
    const uint32_t t = tile_of_block(blockIdx.x - lead, tile_total);
    const SceneConst c = sc.scene_of_tile(t);
    if (!c.use_tiles) return;
    const TileRegion rg = tile_region(c, t - c.tile_base);
    const int emax = min(kE, (int)c.tile_emax);
    const int total = (int)region_runs<kB>(c, b, rg, L.s_off, L.s_beg, L.s_wave);
    if (total == 0) return;
    // the density pass's plan: a region that does not fit at once was handed over
    if (total > 65535 || plan_groups<kB>(rg, total, emax, true, L) != 1u) return;
    stage_group<kB, kOwnMax, true>(b, t, 0u, rg.cells, 0, total, L, s_own_e, s_own_g, [&](int k, uint32_t g) {
        s_a[k] = b.pred_pos[g];
        s_b[k] = b.vel_sorted[g];
    });
    const uint32_t owned = L.s_count;
    if (owned > (uint32_t)kOwnMax) return;  // handed over by the density pass
    const uint16_t *__restrict__ nl16 = b.nlist16();
    for (uint32_t i = tid; i < owned; i += kB) {
        const uint32_t j = s_own_g[i];
        const uint32_t cnt = b.ncount[j];
        const int eo = s_own_e[i];
        const float4 A = s_a[eo], V = s_b[eo];
        const float4 s = make_float4(A.x, A.y, A.z, 1.0f);
        const float4 P = b.pos_sorted[j];
        const float2 myd = make_float2(V.w, A.w);
        const ForceSelf me = force_self(c, j, s, P, V, myd);
        float3 Fp = make_float3(0.0f, 0.0f, 0.0f), Fv = make_float3(0.0f, 0.0f, 0.0f);
        if (cnt != kListInvalid) {
            // the recorded neighbours, in the walk's order: eight indices loaded together,
            // then their terms (the loads' latency is paid once per eight)
            for (uint32_t k0 = 0; k0 < cnt; k0 += 8u) {
                uint32_t ix[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) ix[u] = nl16[(size_t)min(k0 + (uint32_t)u, cnt - 1u) * n + j];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (k0 + (uint32_t)u >= cnt) break;
                    const float4 a = s_a[ix[u]], v = s_b[ix[u]];
                    force_pair<kF>(c, me, make_float3(a.x, a.y, a.z), make_float3(v.x, v.y, v.z),
                                   make_float2(v.w, a.w), Fp, Fv);
                }
            }
        } else {
            // a list that did not fit: the 27 cells from the staged region
            // (applyPressure...comp:288-325).  The particle's own copies need no slot test: at
            // distance 0 the term is skipped (d < 1e-7).
            const int3 c0 = force_cell(c, s);
            const uint32_t cx = (uint32_t)(c0.x - rg.ox), cy = (uint32_t)(c0.y - rg.oy);
            const uint32_t cz = (uint32_t)(c0.z - rg.oz);
            for (uint32_t r = 0; r < 9u; ++r) {
                const uint32_t dz = r / 3u, dy = r % 3u;
                const uint32_t rc0 = ((cz - 1u + dz) * rg.ry + (cy - 1u + dy)) * rg.rx + (cx - 1u);
                const int o3 = L.s_off[rc0 + 3u];
                for (int e = L.s_off[rc0]; e < o3; ++e) {
                    const float4 a = s_a[e];
                    if (dist2(a, s) > c.R2) continue;
                    const float4 v = s_b[e];
                    force_pair<kF>(c, me, make_float3(a.x, a.y, a.z), make_float3(v.x, v.y, v.z),
                                   make_float2(v.w, a.w), Fp, Fv);
                }
            }
        }
        force_store(c, b, j, P, V, myd.x, Fp, Fv);
    }
}

// The particles of tiles handed over by the tiled density pass, one lane each, walking
// their cells through global memory after each tiled pass (recording slot lists when the
// force pass reads them).
constexpr int kStrayBlock = 256;

template <class Scenes, bool kF, bool kS>
__global__ __launch_bounds__(kStrayBlock) void k_density_stray(Scenes sc, const ChainBuffers b, uint32_t n) {
    __shared__ uint32_t s_list[kS ? kListLds * kStrayBlock : 1];
    const uint32_t count = b.counters[7];
    for (uint32_t i = blockIdx.x * kStrayBlock + threadIdx.x; i < count; i += gridDim.x * kStrayBlock) {
        const uint32_t j = b.stray[n - 1u - i];
        if (kS) density_listed<kF>(sc.scene(j), b, n, j, s_list + threadIdx.x, kStrayBlock);
        else density_global<kF>(sc.scene(j), b, j);
    }
}

template <class Scenes, bool kF>
__global__ __launch_bounds__(kStrayBlock) void k_force_stray(Scenes sc, const ChainBuffers b, uint32_t n) {
    __shared__ uint32_t s_q[kQueue * kStrayBlock];
    const uint32_t count = b.counters[7];
    for (uint32_t i = blockIdx.x * kStrayBlock + threadIdx.x; i < count; i += gridDim.x * kStrayBlock) {
        const uint32_t j = b.stray[n - 1u - i];
        const SceneConst c = sc.scene(j);
        const float4 P = b.pos_sorted[j];
        const float4 V = b.vel_sorted[j];
        const float2 myd = b.dens[j];
        const ForceSelf me = force_self(c, j, b.pred_pos[j], P, V, myd);
        const ForceTerms ft = force_global<kF>(c, b, me, s_q + threadIdx.x, kStrayBlock);
        force_store(c, b, j, P, V, myd.x, ft.Fp, ft.Fv);
    }
}

}  // namespace

// Per-lane force over slot lists (lane configuration 3: 4 candidates per scan step, 128
// lanes; 10 its tolerance variant) outruns the tiled force pass wherever both were measured.
const TileConfig kTileConfigs[kNumTileConfigs] = {
    {256, 1280, false, 3}, {128, 768, false, 3}, {256, 1280, false, -1}, {128, 768, false, -1},
    {256, 1280, true, 10},
};

uint32_t scene_tiles(const SceneConst &c) { return c.use_tiles ? c.ntx * c.nty * c.ntz : 0u; }

bool tile_config_tuned(int cfg, bool tolerance) {
    return cfg >= 0 && cfg < kNumTileConfigs && kTileConfigs[cfg].force_lane >= 0 &&
           (tolerance || !kTileConfigs[cfg].tolerance);
}

void tile_dims_for(SceneConst &c, double ppc, int cfg) {
    const TileConfig &tc = kTileConfigs[cfg];
    if (!(ppc > 0.0)) ppc = 1.0;
    // owned particles per tile about 0.75 of the workgroup (on the 1M fill, a tile owning
    // more than a lane each walks twice; 0.75 measured best against 0.85, 0.65 and 0.55),
    // staged entries within 0.8 of the LDS capacity, region cells within kTileRegionMax
    const double own_target = 0.75 * tc.block;
    // a region cell stages its key's whole run, colliding cells' particles included: about
    // one more occupied cell per T / (occupied cells) keys
    const double occupied = std::min<double>((double)(c.p_end - c.p_begin), (double)c.gx * c.gy * c.gz);
    const double collide = 1.0 + 1.3 * occupied / ((double)c.mask + 1.0);
    const double stage_cap = 0.8 * tc.entries / collide;
    // (th, td) in order of preference, the first that leaves tw >= 4: a taller cross-section
    // stages fewer halo entries per owned particle (3 x 3 measured best on the 1M fill)
    static const uint32_t kShapes[4][2] = {{3, 3}, {2, 2}, {2, 1}, {1, 1}};
    uint32_t tw = 1, th = 1, td = 1;
    for (const auto &sh : kShapes) {
        th = sh[0];
        td = sh[1];
        double w = own_target / (ppc * th * td);
        if (w > 64.0) w = 64.0;
        tw = w < 1.0 ? 1u : (uint32_t)w;
        while (tw > 1u && ((double)(tw + 2u) * (th + 2u) * (td + 2u) * ppc > stage_cap ||
                           (tw + 2u) * (th + 2u) * (td + 2u) > kTileRegionMax))
            --tw;
        if (tw >= 4u) break;
    }
    if ((uint32_t)c.gx < tw) tw = (uint32_t)c.gx;
    if ((uint32_t)c.gy < th) th = (uint32_t)c.gy;
    if ((uint32_t)c.gz < td) td = (uint32_t)c.gz;
    // the kernels' LDS tables hold at most kGroupMax rows and kTileRegionMax cells
    if (tw > 64u) tw = 64u;
    c.tw = tw;
    c.th = th;
    c.td = td;
    // staged entries beyond which a tile's region streams in row groups: the LDS capacity
    c.tile_emax = (uint32_t)tc.entries;
    c.ntx = ((uint32_t)c.gx + tw - 1u) / tw;
    c.nty = ((uint32_t)c.gy + th - 1u) / th;
    c.ntz = ((uint32_t)c.gz + td - 1u) / td;
}

namespace {

// Leading stray workgroups of a tiled launch (stray_blocks counts kStrayBlock lanes each;
// a leading workgroup walks with kGlobalLanes of its lanes), a multiple of 8 so that the
// tiles' chunks keep their XCDs.
template <int kB>
uint32_t lead_blocks(uint32_t stray_blocks) {
    return (stray_blocks * (uint32_t)(kStrayBlock / kGlobalLanes<kB>) + 7u) / 8u * 8u;
}

template <class Scenes, int kB, int kE, bool kF, bool kS>
hipError_t launch_density_pair(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                               uint32_t tile_total, uint32_t stray_blocks) {
    const uint32_t lead = lead_blocks<kB>(stray_blocks);
    hipLaunchKernelGGL((k_density_tile<Scenes, kB, kE, kF, kS>), dim3(lead + tile_total), dim3(kB), 0, stream, sc, b,
                       n, lead, tile_total);
    // tiles handed over to the stray list: only without slot lists (with them, a tile's
    // workgroup walks its own particles)
    if (!kS) {
        if (hipPeekAtLastError() != hipSuccess) return hipGetLastError();
        hipLaunchKernelGGL((k_density_stray<Scenes, kF, kS>), dim3(stray_blocks), dim3(kStrayBlock), 0, stream, sc, b,
                           n);
    }
    return hipGetLastError();
}

template <class Scenes, int kB, int kE>
hipError_t launch_force_pair(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                             uint32_t tile_total, uint32_t stray_blocks) {
    const uint32_t lead = lead_blocks<kB>(stray_blocks);
    hipLaunchKernelGGL((k_force_tile<Scenes, kB, kE, false>), dim3(lead + tile_total), dim3(kB), 0, stream, sc, b, n,
                       lead, tile_total);
    if (hipPeekAtLastError() != hipSuccess) return hipGetLastError();
    hipLaunchKernelGGL((k_force_stray<Scenes, false>), dim3(stray_blocks), dim3(kStrayBlock), 0, stream, sc, b, n);
    return hipGetLastError();
}

}  // namespace

template <class Scenes>
hipError_t launch_density_tile(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                               uint32_t tile_total, uint32_t stray_blocks, int cfg) {
    constexpr bool kOne = std::is_same_v<Scenes, OneScene>;
    switch (cfg) {
        case 1: return launch_density_pair<Scenes, 128, 768, false, true>(stream, sc, b, n, tile_total, stray_blocks);
        case 2: return launch_density_pair<Scenes, 256, 1280, false, false>(stream, sc, b, n, tile_total, stray_blocks);
        case 3: return launch_density_pair<Scenes, 128, 768, false, false>(stream, sc, b, n, tile_total, stray_blocks);
        case 4: return launch_density_pair<Scenes, 256, 1280, kOne, true>(stream, sc, b, n, tile_total, stray_blocks);
        default: return launch_density_pair<Scenes, 256, 1280, false, true>(stream, sc, b, n, tile_total, stray_blocks);
    }
}

template <class Scenes>
hipError_t launch_force_tile(hipStream_t stream, const Scenes &sc, const ChainBuffers &b, uint32_t n,
                             uint32_t tile_total, uint32_t stray_blocks, int cfg) {
    if (kTileConfigs[cfg].force_lane >= 0) return launch_force(stream, sc, b, n, kTileConfigs[cfg].force_lane);
    if (kTileConfigs[cfg].block == 128)
        return launch_force_pair<Scenes, 128, 768>(stream, sc, b, n, tile_total, stray_blocks);
    return launch_force_pair<Scenes, 256, 1280>(stream, sc, b, n, tile_total, stray_blocks);
}

#define SPH_TILE_INSTANTIATE(S)                                                                         \
    template hipError_t launch_density_tile<S>(hipStream_t, const S &, const ChainBuffers &, uint32_t, \
                                               uint32_t, uint32_t, int);                               \
    template hipError_t launch_force_tile<S>(hipStream_t, const S &, const ChainBuffers &, uint32_t,   \
                                             uint32_t, uint32_t, int);

SPH_TILE_INSTANTIATE(OneScene)
SPH_TILE_INSTANTIATE(ManyScenes)

}  // namespace sph
