// Per-run neighbour accumulation shared by every density / force kernel variant, so all of
// them evaluate the same float expressions in the same order (bit-identical results).
#pragma once

#include "sph_device.hpp"
#include <type_traits>

namespace sph {

// Table cell of cell (x, y, z): cellCoordToKey (updateCellEntries.comp:59-61) offset by
// the scene's first table cell.  Negative coordinates wrap through uint32.
__device__ __forceinline__ uint32_t key_of_cell(const SceneConst &c, int32_t x, int32_t y, int32_t z) {
    return c.table_base + (cell_hash((uint32_t)x, (uint32_t)y, (uint32_t)z) & c.mask);
}

// calculateDensities.comp:133-152
__device__ __forceinline__ float2 density_finish(const SceneConst &c, float M, float Mn) {
    const float density = M / c.volume;
    const float near_density = Mn * c.near_vol_inv;
    return make_float2(fmaxf(0.01f, density), fmaxf(0.01f, near_density));
}

// Square root and division of the neighbour terms.  The exact path (kFast = false) uses the
// correctly rounded operations the reference's float expressions specify; the tolerance
// path uses the hardware's approximate square root and reciprocal (about 1 ulp each), a
// multiply in place of each division: results then agree with the reference to ~1e-6
// relative instead of bit for bit.
template <bool kFast>
__device__ __forceinline__ float pair_sqrt(float x) {
    if constexpr (kFast) return __builtin_amdgcn_sqrtf(x);
    else return sqrtf(x);
}
template <bool kFast>
__device__ __forceinline__ float pair_div(float a, float b) {
    if constexpr (kFast) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}

struct ForceSelf {
    float4 s;            // predictedPositions[j]
    float3 u;            // predictedVelocities[j]
    float my_pressure;   // convertDensityToPressure(densities[j])
    float near_pressure; // nearDensities[j] * nearPressureMultiplier
    uint32_t j;          // slot of the particle
};

// ---------------------------------------------------------------------------------------
// Neighbour walks.
//
// One lane walks one particle's 27 cells in the reference's (dz, dy, dx) order and each
// run in ascending slot order.  The cell loop is the same for every lane of a wave, so
// the 27 run lookups are issued together; the runs themselves differ in length per lane.
// For the force pass, a loop that did all of a candidate's work would make every lane
// wait for the longest run and execute the expensive in-range arithmetic whenever any lane
// needs it, so each run is only scanned with the distance test, and candidates that pass
// are appended to a per-lane queue in LDS.  Queues are drained, in the order they were filled, when a lane's
// queue is full, after a cell once some lane holds kFlushAt entries, and at the end: every
// particle's sums keep the reference's order exactly, while the expensive arithmetic runs
// with most lanes of the wave active.

constexpr int kQueue = 12;
constexpr int kFlushAt = 6;
constexpr uint32_t kGlobalBit = 0x80000000u;  // queue entry names a global slot

// Edge of a brick in cells and the halo staged around it.
constexpr int kH1 = kBrick + 2;
constexpr int kHalo = kH1 * kH1 * kH1;

// Per halo cell of a brick: (LDS offset of the staged run or -1 when it is read from
// global memory, run length), the run's first slot and the cell's full hash.
struct Halo {
    int2 oc[kHalo];
    int gst[kHalo];
    uint32_t hash[kHalo];
    uint32_t staged;  // staged entries [0, staged)
};

// A run: entries [base, base + cnt) of the LDS copy or of the global arrays.
struct RunRef {
    int base;
    int cnt;
    int first;    // slot of the run's first particle
    bool global;
};

// Finds cell runs: through the brick's halo when the cell lies in it (kHaloLds), else
// through the global start/end table.
template <bool kHaloLds>
struct CellRuns {
    uint32_t table_base, mask;  // the scene's first table cell and key mask
    const int2 *__restrict__ table;
    const Halo *H;
    int ox, oy, oz;
    __device__ __forceinline__ RunRef get(int cx, int cy, int cz) const {
        if (kHaloLds) {
            const int lx = cx - ox, ly = cy - oy, lz = cz - oz;
            if ((unsigned)lx < (unsigned)kH1 && (unsigned)ly < (unsigned)kH1 &&
                (unsigned)lz < (unsigned)kH1) {
                const int h = lx + kH1 * (ly + kH1 * lz);
                const int2 oc = H->oc[h];
                const int g = H->gst[h];
                return RunRef{oc.x >= 0 ? oc.x : g, oc.y, g, oc.x < 0};
            }
        }
        const int2 se = table[table_base + (cell_hash((uint32_t)cx, (uint32_t)cy, (uint32_t)cz) & mask)];
        if (se.x == -1) return RunRef{0, 0, 0, true};
        return RunRef{se.x, se.y - se.x, se.x, true};
    }
};

// Whether cells cx - 1 .. cx + 1 of a row all lie in a halo-grid row of hx >= 2 entries
// (entry l holds cell l - 1), i.e. 0 <= cx and cx + 2 < hx, as one unsigned compare.  The
// form lx + 2 < hx with lx = (uint32)cx wraps for cx = -1 and -2 (the cell of a point
// outside the box) and would admit an index 2^32 - 1 past the row.
__host__ __device__ constexpr bool row3_in_grid(int32_t cx, uint32_t hx) {
    return (uint32_t)cx < hx - 2u;
}

// Finds cell runs in the cell-ordered layout: through the scene's halo grid when the cell
// lies in it (hz = 0: the scene has none), else through the key's cell-ordered run.
struct GridRuns {
    const int2 *__restrict__ grid;
    uint32_t grid_base, hx, hy, hz;
    uint32_t table_base, mask;
    const int2 *__restrict__ crun;
    // Raw (start, end) in cell order; start < 0: no run.  Loads are left unconsumed so that
    // several lookups can be in flight before any is used.
    __device__ __forceinline__ int2 slow_raw(int cx, int cy, int cz) const {
        return crun[table_base + (cell_hash((uint32_t)cx, (uint32_t)cy, (uint32_t)cz) & mask)];
    }
    __device__ __forceinline__ int2 raw(int cx, int cy, int cz) const {
        const uint32_t lx = (uint32_t)(cx + 1), ly = (uint32_t)(cy + 1), lz = (uint32_t)(cz + 1);
        if (lx < hx && ly < hy && lz < hz) return grid[grid_base + lx + hx * (ly + hy * lz)];
        return slow_raw(cx, cy, cz);
    }
    // The three cells (cx - 1 .. cx + 1, cy, cz) of a row: consecutive grid entries.
    __device__ __forceinline__ void row3(int cx, int cy, int cz, int2 out[3]) const {
        const uint32_t lx = (uint32_t)cx, ly = (uint32_t)(cy + 1), lz = (uint32_t)(cz + 1);
        if (row3_in_grid(cx, hx) && ly < hy && lz < hz) {
            const int2 *g = grid + grid_base + lx + hx * (ly + hy * lz);
            out[0] = g[0];
            out[1] = g[1];
            out[2] = g[2];
        } else {
            for (int d = 0; d < 3; ++d) out[d] = raw(cx + d - 1, cy, cz);
        }
    }
    static __device__ __forceinline__ RunRef ref(const int2 &se) {
        const bool empty = se.x < 0;
        const int base = empty ? 0 : se.x;
        return RunRef{base, empty ? 0 : se.y - se.x, base, true};
    }
    __device__ __forceinline__ RunRef get(int cx, int cy, int cz) const { return ref(raw(cx, cy, cz)); }
};

// The density pass's 27 cells (calculateDensities.comp:114-118): sample points
// s + R*(dx, dy, dz); a sample point outside the box skips its cell.
template <class Runs>
struct DensityCellsOf {
    Runs runs;
    AxisCells ax, ay, az;
    __device__ __forceinline__ RunRef operator()(int dz, int dy, int dx) const {
        if (!(ax.in(dx) && ay.in(dy) && az.in(dz))) return RunRef{0, 0, 0, true};
        return runs.get(ax.c(dx), ay.c(dy), az.c(dz));
    }
    // The three cells of row (dz, dy); with the grid lookup, consecutive cells are loaded
    // together and non-member cells masked afterwards.
    __device__ __forceinline__ void row(int dz, int dy, RunRef r[3]) const {
        if constexpr (std::is_same_v<Runs, GridRuns>) {
            if (ax.lo_off == 0xFFFFFFFFu && ax.hi_off == 1u) {
                int2 raw[3];
                runs.row3((int32_t)ax.mid, ay.c(dy), az.c(dz), raw);
                const bool row_in = ay.in(dy) && az.in(dz);
                for (int dx = 0; dx < 3; ++dx) {
                    const RunRef q = GridRuns::ref(raw[dx]);
                    r[dx] = (row_in && ax.in(dx)) ? q : RunRef{0, 0, 0, true};
                }
                return;
            }
        }
        for (int dx = 0; dx < 3; ++dx) r[dx] = (*this)(dz, dy, dx);
    }
};

// The force pass's 27 cells (applyPressure...comp:283-292): the sample point's cell plus
// integer offsets, no bounds skip.
template <class Runs>
struct ForceCellsOf {
    Runs runs;
    int3 c0;
    __device__ __forceinline__ RunRef operator()(int dz, int dy, int dx) const {
        return runs.get(c0.x + dx - 1, c0.y + dy - 1, c0.z + dz - 1);
    }
    __device__ __forceinline__ void row(int dz, int dy, RunRef r[3]) const {
        if constexpr (std::is_same_v<Runs, GridRuns>) {
            int2 raw[3];
            runs.row3(c0.x, c0.y + dy - 1, c0.z + dz - 1, raw);
            for (int dx = 0; dx < 3; ++dx) r[dx] = GridRuns::ref(raw[dx]);
        } else {
            for (int dx = 0; dx < 3; ++dx) r[dx] = (*this)(dz, dy, dx);
        }
    }
};

template <bool kHaloLds>
using DensityCells = DensityCellsOf<CellRuns<kHaloLds>>;
template <bool kHaloLds>
using ForceCells = ForceCellsOf<CellRuns<kHaloLds>>;

// One density candidate at distance^2 d2 (calculateDensities.comp:123-127).  d2 > R2_pre
// implies sqrt(d2) >= R: both kernels are 0 and M + m*0 == M for M >= +0, so such a
// candidate changes nothing and its sqrt is skipped.
// The terms of a candidate that passed the d2 <= R2_pre pre-filter.
template <bool kFast = false>
__device__ __forceinline__ void density_terms(const SceneConst &c, float d2, float &M, float &Mn) {
    const float dist = pair_sqrt<kFast>(d2);
    if (dist >= c.R) return;
    const float t = c.R - dist;
    M = M + c.m * (t * t);
    Mn = Mn + c.m * (t * t * t);
}

// density_terms without branches, for candidates that `take`; the others add +0, and
// M + (+0) == M, Mn + (+-0) == Mn (both sums are >= +0): exactly the reference's skipped
// terms.  For queued candidates, nearly all in range.
template <bool kFast = false>
__device__ __forceinline__ void density_terms_masked(const SceneConst &c, float d2, bool take, float &M, float &Mn) {
    const float dist = pair_sqrt<kFast>(d2);
    const float t = c.R - dist;
    const float tt = take && dist < c.R ? t * t : 0.0f;
    M = M + c.m * tt;
    Mn = Mn + c.m * (tt * t);
}

template <bool kFast = false>
__device__ __forceinline__ void density_accumulate(const SceneConst &c, float d2, float &M, float &Mn) {
    if (d2 > c.R2_pre) return;
    const float dist = pair_sqrt<kFast>(d2);
    if (dist >= c.R) return;
    const float t = c.R - dist;
    M = M + c.m * (t * t);
    Mn = Mn + c.m * (t * t * t);
}

__device__ __forceinline__ float dist2(const float4 &p, const float4 &s) {
    const float ex = p.x - s.x, ey = p.y - s.y, ez = p.z - s.z;
    return ex * ex + ey * ey + ez * ez;
}

// calculateDensities.comp:114-134 for sample point s.  With kLds false every run is read
// from global memory and s_pos is unused.  kU candidates of a run are loaded before any
// is used (more loads in flight, fewer loop branches); they are still accumulated one by
// one in run order.  The density's in-range arithmetic is short, so it is done in the
// scan itself (no queue).
template <bool kLds, int kU, class Cells>
__device__ __forceinline__ float2 density_walk(const SceneConst &c, const float4 &s,
                                               const float4 *__restrict__ pred_pos,
                                               const float4 *s_pos, const Cells &cells) {
    float M = 0.0f, Mn = 0.0f;
    for (int dz = 0; dz < 3; ++dz)
        for (int dy = 0; dy < 3; ++dy)
            for (int dx = 0; dx < 3; ++dx) {
                const RunRef r = cells(dz, dy, dx);
                // Separate loops per source keep global and LDS loads in their own
                // address spaces (no flat loads).
                auto scan = [&](const float4 *src) {
                    int k = 0;
                    for (; k + kU <= r.cnt; k += kU) {
                        float4 p[kU];
#pragma unroll
                        for (int u = 0; u < kU; ++u) p[u] = src[k + u];
#pragma unroll
                        for (int u = 0; u < kU; ++u) density_accumulate(c, dist2(p[u], s), M, Mn);
                    }
                    for (; k < r.cnt; ++k) density_accumulate(c, dist2(src[k], s), M, Mn);
                };
                if constexpr (!kLds) {
                    scan(pred_pos + r.base);
                } else {
                    if (r.global) scan(pred_pos + r.base);
                    else scan(s_pos + r.base);
                }
            }
    return density_finish(c, M, Mn);
}

// Full arithmetic of one force candidate (applyPressureAndViscosityForce.comp:303-321) at
// predicted position pi with velocity vi and (density, near density) di.
template <bool kFast = false>
__device__ __forceinline__ void force_pair(const SceneConst &c, const ForceSelf &me, const float3 &pi,
                                           const float3 &vi, const float2 &di, float3 &Fp, float3 &Fv) {
    const float ex = pi.x - me.s.x, ey = pi.y - me.s.y, ez = pi.z - me.s.z;
    const float d2 = ex * ex + ey * ey + ez * ez;
    const float dist = pair_sqrt<kFast>(d2);
    if (fabsf(dist) < 1e-7f) return;
    const float dirx = pair_div<kFast>(ex, dist), diry = pair_div<kFast>(ey, dist), dirz = pair_div<kFast>(ez, dist);
    const float slope = dist >= c.R ? 0.0f : (dist - c.R) * c.slope_scale;
    const float shared = ((di.x - c.rho0) * c.k + me.my_pressure) / 2.0f;
    const float a = pair_div<kFast>(shared * slope * c.m, di.x);
    Fp.x = Fp.x + a * dirx;
    Fp.y = Fp.y + a * diry;
    Fp.z = Fp.z + a * dirz;
    const float infl = dist >= c.R ? 0.0f : (c.R - dist) * (c.R - dist);
    const float b = pair_div<kFast>(c.m * infl, di.x);
    Fv.x = Fv.x + b * (vi.x - me.u.x);
    Fv.y = Fv.y + b * (vi.y - me.u.y);
    Fv.z = Fv.z + b * (vi.z - me.u.z);
    const float shared_near = (me.near_pressure + di.y * c.kn) / 2.0f;
    const float x = c.R - dist;
    const float nslope = dist >= c.R ? 0.0f : x * x * c.near_slope_scale;
    const float cn = pair_div<kFast>(shared_near * nslope * c.m, di.y);
    Fp.x = Fp.x + cn * dirx;
    Fp.y = Fp.y + cn * diry;
    Fp.z = Fp.z + cn * dirz;
}

// One force candidate by queue entry e: a global slot (kGlobalBit) or, on the brick path,
// an LDS entry packed as s_a = (predicted position xyz, near density) and s_b = (velocity
// xyz, density).
template <bool kLds, bool kFast = false>
__device__ __forceinline__ void force_term(const SceneConst &c, const ForceSelf &me, uint32_t e,
                                           const float4 *__restrict__ pred_pos,
                                           const float4 *__restrict__ vel_sorted,
                                           const float2 *__restrict__ dens, const float4 *s_a,
                                           const float4 *s_b, float3 &Fp, float3 &Fv) {
    if (!kLds || (e & kGlobalBit)) {
        const uint32_t g = e & ~kGlobalBit;
        const float4 a = pred_pos[g], b = vel_sorted[g];
        force_pair<kFast>(c, me, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), dens[g], Fp, Fv);
    } else {
        const float4 a = s_a[e], b = s_b[e];
        force_pair<kFast>(c, me, make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), make_float2(b.w, a.w), Fp, Fv);
    }
}

// The force terms of a recorded neighbour list (entry i at list[i * n + j]), in order.
// kU entries and their neighbours' data are loaded before any term is evaluated (the
// pass waits on these dependent loads, not on the arithmetic).  The density pass that
// recorded the list left each particle's density in vel_sorted[].w and its near density
// in pred_pos[].w, so a neighbour is two 16-byte loads.
template <int kU, bool kFast = false>
__device__ __forceinline__ void force_list(const SceneConst &c, const ForceSelf &me, uint32_t cnt,
                                           const uint32_t *__restrict__ list, uint32_t n,
                                           const float4 *__restrict__ pred_pos,
                                           const float4 *__restrict__ vel_sorted, float3 &Fp, float3 &Fv) {
    for (uint32_t i = 0; i < cnt; i += kU) {
        uint32_t g[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) g[u] = list[(size_t)min(i + (uint32_t)u, cnt - 1u) * n + me.j];
        float4 a[kU], v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            a[u] = pred_pos[g[u]];
            v[u] = vel_sorted[g[u]];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (i + (uint32_t)u < cnt)
                force_pair<kFast>(c, me, make_float3(a[u].x, a[u].y, a[u].z), make_float3(v[u].x, v[u].y, v[u].z),
                           make_float2(v[u].w, a[u].w), Fp, Fv);
    }
}

// applyPressureAndViscosityForce.comp:288-325: accumulates Fp and Fv for `me`.  The scan
// loads kU candidates of a run before testing them in run order.
template <bool kLds, int kU, class Cells>
__device__ __forceinline__ void force_walk(const SceneConst &c, const ForceSelf &me,
                                           const float4 *__restrict__ pred_pos,
                                           const float4 *__restrict__ vel_sorted,
                                           const float2 *__restrict__ dens,
                                           const float4 *s_a, const float4 *s_b,
                                           const Cells &cells, uint32_t *q, int stride,
                                           float3 &Fp, float3 &Fv) {
    int qc = 0;
    auto drain = [&]() {
        for (int i = 0; i < qc; ++i)
            force_term<kLds>(c, me, q[i * stride], pred_pos, vel_sorted, dens, s_a, s_b, Fp, Fv);
        qc = 0;
    };
    auto consider = [&](const float4 &p, int k, const RunRef &r, bool global) {
        if ((uint32_t)(r.first + k) == me.j) return;
        if (dist2(p, me.s) > c.R2) return;
        const uint32_t idx = (uint32_t)(r.base + k);
        q[qc * stride] = global ? (kGlobalBit | idx) : idx;
        if (++qc == kQueue) drain();
    };
    for (int dz = 0; dz < 3; ++dz)
        for (int dy = 0; dy < 3; ++dy)
            for (int dx = 0; dx < 3; ++dx) {
                const RunRef r = cells(dz, dy, dx);
                const bool global = !kLds || r.global;
                auto scan = [&](const float4 *src) {
                    int k = 0;
                    for (; k + kU <= r.cnt; k += kU) {
                        float4 p[kU];
#pragma unroll
                        for (int u = 0; u < kU; ++u) p[u] = src[k + u];
#pragma unroll
                        for (int u = 0; u < kU; ++u) consider(p[u], k + u, r, global);
                    }
                    for (; k < r.cnt; ++k) consider(src[k], k, r, global);
                };
                if constexpr (!kLds) {
                    scan(pred_pos + r.base);
                } else {
                    if (global) scan(pred_pos + r.base);
                    else scan(s_a + r.base);
                }
                if (__ballot(qc >= kFlushAt)) drain();
            }
    drain();
}

// ---------------------------------------------------------------------------------------
// Walks of the per-lane kernels (runs in global memory).  The walks are bound by the
// latency of dependent loads, so they are arranged for loads in flight: the three run
// lookups of a (dz, dy) row are issued together, and every scan step loads kU candidates
// (indices clamped to the run, extra lanes masked) so that a run of up to kU candidates
// costs one round trip.  Sums still take the candidates one by one in run order.

// A candidate's predicted position.  (Loading its three dwords only, 12 bytes instead of
// 16 per lane, measured no faster: the walks are not bound by the bytes they fetch.)
__device__ __forceinline__ float4 load_xyz(const float4 *p) { return *p; }

// The kU candidates of a run loaded together at run offset k: only lanes whose run still
// has candidate u issue its load (the others would only repeat a load, and every load costs
// the texture path its lanes' bytes); candidate 0 always exists inside the loop.
template <int kU>
__device__ __forceinline__ void load_candidates(const float4 *src, int k, int cnt, float4 (&p)[kU]) {
#pragma unroll
    for (int u = 0; u < kU; ++u)
        p[u] = (u == 0 || k + u < cnt) ? load_xyz(src + k + u) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

template <int kU, bool kFast = false, class Cells>
__device__ __forceinline__ float2 density_walk_rows(const SceneConst &c, const float4 &s,
                                                    const float4 *__restrict__ pred_pos,
                                                    const Cells &cells) {
    float M = 0.0f, Mn = 0.0f;
    for (int dz = 0; dz < 3; ++dz)
        for (int dy = 0; dy < 3; ++dy) {
            RunRef r[3];
            cells.row(dz, dy, r);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const float4 *src = pred_pos + r[dx].base;
                const int cnt = r[dx].cnt;
                for (int k = 0; k < cnt; k += kU) {
                    float4 p[kU];
                    load_candidates<kU>(src, k, cnt, p);
#pragma unroll
                    for (int u = 0; u < kU; ++u)
                        if (k + u < cnt) density_accumulate<kFast>(c, dist2(p[u], s), M, Mn);
                }
            }
        }
    return density_finish(c, M, Mn);
}

template <int kU, bool kFast = false, class Cells>
__device__ __forceinline__ void force_walk_rows(const SceneConst &c, const ForceSelf &me,
                                                const float4 *__restrict__ pred_pos,
                                                const float4 *__restrict__ vel_sorted,
                                                const float2 *__restrict__ dens, const Cells &cells,
                                                uint32_t *q, int stride, float3 &Fp, float3 &Fv) {
    int qc = 0;
    auto drain = [&]() {
        for (int i = 0; i < qc; ++i)
            force_term<false, kFast>(c, me, q[i * stride], pred_pos, vel_sorted, dens, nullptr, nullptr, Fp, Fv);
        qc = 0;
    };
    for (int dz = 0; dz < 3; ++dz)
        for (int dy = 0; dy < 3; ++dy) {
            RunRef r[3];
            cells.row(dz, dy, r);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const float4 *src = pred_pos + r[dx].base;
                const int cnt = r[dx].cnt;
                for (int k = 0; k < cnt; k += kU) {
                    float4 p[kU];
                    load_candidates<kU>(src, k, cnt, p);
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if (k + u >= cnt) continue;
                        if ((uint32_t)(r[dx].first + k + u) == me.j) continue;
                        if (dist2(p[u], me.s) > c.R2) continue;
                        q[qc * stride] = kGlobalBit | (uint32_t)(r[dx].base + k + u);
                        if (++qc == kQueue) drain();
                    }
                }
                if (__ballot(qc >= kFlushAt)) drain();
            }
        }
    drain();
}

// Self terms of the force pass for slot j (applyPressureAndViscosityForce.comp:276-286).
__device__ __forceinline__ ForceSelf force_self(const SceneConst &c, uint32_t j, const float4 &s,
                                                const float4 &P, const float4 &V,
                                                const float2 &myd) {
    ForceSelf me;
    me.s = s;
    float3 ps = make_float3(P.x, P.y, P.z);
    me.u = make_float3(V.x, V.y, V.z);
    kick_drift<false>(c, ps, me.u);
    me.my_pressure = (myd.x - c.rho0) * c.k;
    me.near_pressure = myd.y * c.kn;
    me.j = j;
    return me;
}

// Cell coordinate of the force pass's sample point (samplePointToCellCoord,
// applyPressure...comp:224-228): -1 on every axis when outside the box.
__device__ __forceinline__ int3 force_cell(const SceneConst &c, const float4 &s) {
    const bool outside = s.x < c.minx || s.y < c.miny || s.z < c.minz || s.x > c.maxx ||
                         s.y > c.maxy || s.z > c.maxz;
    if (outside) return make_int3(-1, -1, -1);
    return make_int3(cell_of(s.x, c.minx, c.R), cell_of(s.y, c.miny, c.R), cell_of(s.z, c.minz, c.R));
}

// velocities[j] += vec4((Fp + Fv + 0) / densities[j] * dt, 1) and updateVelocities.
__device__ __forceinline__ void force_finish(const SceneConst &c, const float4 &P, const float4 &V,
                                             float my_density, const float3 &Fp, const float3 &Fv,
                                             float4 &pos_out, float4 &vel_out, float4 &vel_force) {
    float3 v = make_float3(V.x + (Fp.x + Fv.x + 0.0f) / my_density * c.dt,
                           V.y + (Fp.y + Fv.y + 0.0f) / my_density * c.dt,
                           V.z + (Fp.z + Fv.z + 0.0f) / my_density * c.dt);
    vel_force = make_float4(v.x, v.y, v.z, V.w + 1.0f);
    float3 p = make_float3(P.x, P.y, P.z);
    kick_drift<true>(c, p, v);
    pos_out = make_float4(p.x, p.y, p.z, 1.0f);
    vel_out = make_float4(v.x, v.y, v.z, 1.0f);
}


// ---------------------------------------------------------------------------------------
// Force neighbour lists, recorded by the density pass.
//
// The force pass visits the cells c0 + (dx, dy, dz) of the predicted position's cell c0
// (applyPressure...comp:283-292); the density pass visits the cells of the sample points
// s + R*(dx, dy, dz), skipping sample points outside the box (calculateDensities.comp:
// 114-118).  Per axis the middle sample point is s itself, whose cell is c0, and the other
// two normally fall in c0 -/+ 1.  When that holds on every axis for every sample point
// inside the box, the density's cells are exactly the force cells whose sample points are
// inside, in the same (dz, dy, dx) order, so one walk over the force cells yields both:
// density terms from the member cells, and the force candidates (slot != j, d2 <= R*R) in
// the force pass's visit order.  The first kListLds candidates are appended to a per-lane
// list in LDS and copied out entry-major (entry k of slot j at list[k * n + j]: one
// coalesced row per entry), the few beyond are stored directly; the force pass then reads
// its list instead of scanning 27 cells.  A particle for
// which the cell relation fails (a rounded sample cell that is not c0 -/+ 1, a predicted
// position outside the box), or whose list would exceed the `cap` entries allocated per
// slot (at most kList), gets an invalid count, and the force pass walks its cells; `found`
// returns the candidates the walk recorded (0 when it did not record), so that the kernel
// can report lists that did not fit.
constexpr int kListLds = 20;   // entries staged in LDS (the rest are stored directly)

template <int kU, bool kFast = false>
__device__ __forceinline__ float2 density_with_list(const SceneConst &c, uint32_t j, const float4 &s,
                                                    const float4 *__restrict__ pred_pos,
                                                    const GridRuns &runs, uint32_t *lane_list,
                                                    int stride, uint32_t *__restrict__ nlist,
                                                    uint32_t n, uint32_t cap, uint32_t &count,
                                                    uint32_t &found) {
    const AxisCells ax = sample_axis(s.x, c.minx, c.maxx, c.R);
    const AxisCells ay = sample_axis(s.y, c.miny, c.maxy, c.R);
    const AxisCells az = sample_axis(s.z, c.minz, c.maxz, c.R);
    const int3 c0 = force_cell(c, s);
    auto same = [](const AxisCells &a) {
        return a.in(1) && (!a.in(0) || a.lo_off == 0xFFFFFFFFu) && (!a.in(2) || a.hi_off == 1u);
    };
    if (!(same(ax) && same(ay) && same(az) && (int32_t)ax.mid == c0.x && (int32_t)ay.mid == c0.y &&
          (int32_t)az.mid == c0.z)) {
        count = kListInvalid;
        found = 0;
        return density_walk_rows<kU, kFast>(c, s, pred_pos, DensityCellsOf<GridRuns>{runs, ax, ay, az});
    }
    const float R2 = c.R2;
    float M = 0.0f, Mn = 0.0f;
    uint32_t nl = 0;
    for (int dz = 0; dz < 3; ++dz)
        for (int dy = 0; dy < 3; ++dy) {
            int2 raw[3];
            runs.row3(c0.x, c0.y + dy - 1, c0.z + dz - 1, raw);
            RunRef r[3];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) r[dx] = GridRuns::ref(raw[dx]);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const bool member = ax.in(dx) && ay.in(dy) && az.in(dz);
                const float4 *src = pred_pos + r[dx].base;
                const int cnt = r[dx].cnt;
                for (int k = 0; k < cnt; k += kU) {
                    float4 p[kU];
                    load_candidates<kU>(src, k, cnt, p);
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if (k + u >= cnt) continue;
                        const float d2 = dist2(p[u], s);
                        // one branch for both uses: R2 <= R2_pre, and most candidates fail
                        // (whole waves often: computing the terms without it measured 23 %
                        // slower)
                        if (!(d2 > c.R2_pre)) {
                            if (member) density_terms<kFast>(c, d2, M, Mn);
                            const uint32_t slot = (uint32_t)(r[dx].base + k + u);
                            if (slot != j && !(d2 > R2)) {
                                if (nl < (uint32_t)kListLds) lane_list[nl * stride] = slot;
                                else if (nl < cap) nlist[(size_t)nl * n + j] = slot;
                                ++nl;
                            }
                        }
                    }
                }
            }
        }
    count = nl <= cap ? nl : kListInvalid;
    found = nl;
    return density_finish(c, M, Mn);
}

}  // namespace sph
