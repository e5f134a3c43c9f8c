// Density map for the volumetric renderer: updateDensityMap.comp:114-148, dispatched after
// the substeps (commands.c:418-424) over a voxel grid of edge DENSITY_CELL_SIZE (main.c:14).
//
// One lane per voxel, voxels numbered x-fastest (the image's x, y, z).  Each voxel centre
// walks the 27 cells around its own cell in (dz, dy, dx) order and sums
// m * (R - d)^2 over the particles of each cell's run closer than R, reading the positions
// and the start/end table the last substep left: positions in its sorted order, runs keyed
// by the cells the particles were hashed in at the start of that substep.  As in the
// shader, cells are bounds-checked against the voxel counts, not the cell counts, and the
// corner voxel (M, M, M) is skipped.  Lanes of a wave hold neighbouring voxels, which share
// most of their cells.
//
// After a substep of the per-lane kernels the same runs, in the same order, are also laid
// out in cell order (the force pass stores the new positions there too), where the runs of
// neighbouring cells sit next to each other; the map then reads them through the halo grid
// and the cell-ordered run table instead of the hash-order table.  Either source gives the
// same sums bit for bit.
#include "sph_device.hpp"
#include "sph_map.hpp"
#include "sph_pairs.hpp"

namespace sph {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ bool map_out_of_bounds(int x, int y, int z, const MapGeometry &g) {
    if (x < 0 || y < 0 || z < 0) return true;
    if (x > g.mx || y > g.my || z > g.mz) return true;
    return x == g.mx && y == g.my && z == g.mz;
}

// Runs of the hash-order table: (start, end) of the cell's key, start -1 when empty.
struct TableRuns {
    const int2 *__restrict__ table;
    uint32_t table_base, mask;
    __device__ __forceinline__ int2 raw(int x, int y, int z) const {
        return table[table_base + (cell_hash((uint32_t)x, (uint32_t)y, (uint32_t)z) & mask)];
    }
    __device__ __forceinline__ void row3(int x, int y, int z, int2 out[3]) const {
        for (int d = 0; d < 3; ++d) out[d] = raw(x + d - 1, y, z);
    }
};

// Sums the density of K voxels of one x row (centres cx[k], common cy, cz) over the cells
// of rows oy + dy, oz + dz and columns oxlo - 1 .. oxhi + 1 (oxhi - oxlo <= 1): each run is
// read once, and voxel k takes a candidate only from its own columns ox[k] - 1 .. ox[k] + 1.
// Rows in (dz, dy) order, columns in increasing x and each run in slot order are the
// shader's visit order for every voxel, so each sum is the one-voxel walk's bit for bit.
template <int K, class Runs>
__device__ __forceinline__ void map_walk(const SceneConst &c, const float4 *__restrict__ pos, const Runs &runs,
                                         const MapGeometry &g, int oy, int oz, float cy, float cz,
                                         const float (&cx)[K], const int (&ox)[K], const bool (&live)[K],
                                         int oxlo, int oxhi, float (&dens)[K]) {
    const float R = c.R, R2 = c.R2;
    constexpr int kU = 4;  // candidates of a run loaded together
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = oy + dy, z = oz + dz;
            if (y < 0 || z < 0 || y > g.my || z > g.mz) continue;  // the whole row is out
            int2 se[4];
            runs.row3(oxlo, y, z, se);  // columns oxlo - 1 .. oxlo + 1, loaded together
            se[3] = oxhi > oxlo ? runs.raw(oxhi + 1, y, z) : make_int2(-1, -1);
#pragma unroll
            for (int col = 0; col < 4; ++col) {
                const int x = oxlo - 1 + col;
                if (x > oxhi + 1 || map_out_of_bounds(x, y, z, g)) continue;
                if (se[col].x == -1) continue;
                bool take[K];
#pragma unroll
                for (int k = 0; k < K; ++k) take[k] = live[k] && x >= ox[k] - 1 && x <= ox[k] + 1;
                const float4 *src = pos + se[col].x;
                const int cnt = se[col].y - se[col].x;
                for (int i = 0; i < cnt; i += kU) {
                    float4 p[kU];
                    load_candidates<kU>(src, i, cnt, p);
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if (i + u >= cnt) continue;
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            if (!take[k]) continue;
                            const float ex = cx[k] - p[u].x, ey = cy - p[u].y, ez = cz - p[u].z;
                            const float d2 = ex * ex + ey * ey + ez * ez;
                            if (d2 < R2) {
                                const float dist = sqrtf(d2);
                                const float t = dist >= R ? 0.0f : (R - dist) * (R - dist);
                                dens[k] = dens[k] + c.m * t;
                            }
                        }
                    }
                }
            }
        }
}

// One lane per K consecutive voxels of an x row (updateDensityMap.comp:114-148 per voxel).
// Neighbouring voxels share most of their 27 cells: when the lane's voxels lie inside the
// box and span at most two cell columns, their cells are walked once for all of them;
// otherwise (a voxel centre outside the box, or voxels spread over three columns) each is
// walked alone.
// The positions a per-lane substep wrote back to the hash slots, gathered into cell order
// for the walk over the cell-ordered runs: cell slot j of the scene takes hash slot hslot[j].
__global__ __launch_bounds__(kBlock) void k_gather_cell_positions(const float4 *__restrict__ pos,
                                                                  const uint32_t *__restrict__ hslot,
                                                                  float4 *__restrict__ out, uint32_t begin,
                                                                  uint32_t end) {
    const uint32_t j = begin + blockIdx.x * kBlock + threadIdx.x;
    if (j < end) out[j] = pos[hslot[j]];
}

template <class Runs, int K>
__global__ __launch_bounds__(kBlock) void k_density_map(SceneConst c, const float4 *__restrict__ pos,
                                                       Runs runs, MapGeometry g,
                                                       float *__restrict__ out) {
    const uint64_t per_row = ((uint64_t)g.dx + K - 1) / K;
    const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= per_row * g.dy * g.dz) return;
    const uint64_t row = t / per_row;
    const int vx0 = (int)(t - row * per_row) * K;
    const int vy = (int)(row % g.dy), vz = (int)(row / g.dy);
    const uint64_t v0 = row * g.dx + (uint64_t)vx0;
    const float cy = c.miny + ((float)vy + 0.5f) * g.voxel;
    const float cz = c.minz + ((float)vz + 0.5f) * g.voxel;
    float cx[K], dens[K];
    int ox[K];
    bool live[K], outside[K];
    bool joint = true;
    int oxlo = 0x7FFFFFFF, oxhi = -0x7FFFFFFF;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int vx = vx0 + k;
        live[k] = vx < (int)g.dx && !map_out_of_bounds(vx, vy, vz, g);
        cx[k] = c.minx + ((float)vx + 0.5f) * g.voxel;
        dens[k] = 0.0f;
        // samplePointToCellCoord: -1 on every axis when the centre lies outside the box
        outside[k] = cx[k] < c.minx || cy < c.miny || cz < c.minz || cx[k] > c.maxx || cy > c.maxy ||
                     cz > c.maxz;
        ox[k] = outside[k] ? -1 : cell_of(cx[k], c.minx, c.R);
        if (live[k]) {
            joint = joint && !outside[k];
            oxlo = min(oxlo, ox[k]);
            oxhi = max(oxhi, ox[k]);
        }
    }
    if (oxlo <= oxhi) {  // some voxel to evaluate
        if (joint && oxhi - oxlo <= 1) {
            map_walk<K>(c, pos, runs, g, cell_of(cy, c.miny, c.R), cell_of(cz, c.minz, c.R), cy, cz, cx, ox,
                        live, oxlo, oxhi, dens);
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (!live[k]) continue;
                const float cx1[1] = {cx[k]};
                const int ox1[1] = {ox[k]};
                const bool live1[1] = {true};
                float d1[1] = {0.0f};
                map_walk<1>(c, pos, runs, g, outside[k] ? -1 : cell_of(cy, c.miny, c.R),
                            outside[k] ? -1 : cell_of(cz, c.minz, c.R), cy, cz, cx1, ox1, live1, ox[k], ox[k], d1);
                dens[k] = d1[0];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (vx0 + k >= (int)g.dx) break;
        out[v0 + k] = live[k] ? (dens[k] * 1.0f) / c.volume : 0.0f;
    }
}

}  // namespace

int map_geometry(const sph_params &p, float voxel_size, MapGeometry &g) {
    if (!std::isfinite(voxel_size) || !(voxel_size > 0.0f))
        return fail(SPH_ERR_INVALID_ARG, "voxel_size must be finite and > 0");
    // initializeDensityMap (renderer.c:774-783): extent times the float reciprocal, ceil
    const float inv = 1.0f / voxel_size;
    uint64_t d[3];
    int32_t m[3];
    for (int a = 0; a < 3; ++a) {
        const double e = std::ceil((double)((p.max_xyz[a] - p.min_xyz[a]) * inv));
        // the shader's own count: ceil((max - min) / voxel) (updateDensityMap.comp:95)
        const double s = std::ceil((double)((p.max_xyz[a] - p.min_xyz[a]) / voxel_size));
        if (!(e >= 1.0) || e > 1.0e9 || s > 1.0e9)
            return fail(SPH_ERR_INVALID_ARG, "voxel grid of %g voxels per axis is out of range", e);
        d[a] = (uint64_t)e;
        m[a] = (int32_t)s;
    }
    const uint64_t total = d[0] * d[1] * d[2];
    if (total >= (1ull << 34)) return fail(SPH_ERR_CAPACITY, "density map of %llu voxels is too large",
                                           (unsigned long long)total);
    g.voxel = voxel_size;
    g.dx = (uint32_t)d[0];
    g.dy = (uint32_t)d[1];
    g.dz = (uint32_t)d[2];
    g.mx = m[0];
    g.my = m[1];
    g.mz = m[2];
    g.voxels = total;
    return SPH_OK;
}

hipError_t launch_density_map(hipStream_t stream, const SceneConst &c, const Chain &chain,
                              bool grid_current, const MapGeometry &g, float *out) {
    // voxels per lane: as many as keep their centres within R of each other (two columns)
    const int K = 3.0f * g.voxel <= c.R ? 4 : (g.voxel <= c.R ? 2 : 1);
    const uint64_t lanes = ((uint64_t)g.dx + K - 1) / K * g.dy * g.dz;
    const uint64_t blocks = (lanes + kBlock - 1u) / kBlock;
    const ChainBuffers &b = chain.b;
    auto launch = [&](auto runs, const float4 *p) {
        using R = decltype(runs);
        if (K == 4)
            hipLaunchKernelGGL((k_density_map<R, 4>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, c, p, runs, g, out);
        else if (K == 2)
            hipLaunchKernelGGL((k_density_map<R, 2>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, c, p, runs, g, out);
        else
            hipLaunchKernelGGL((k_density_map<R, 1>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, c, p, runs, g, out);
    };
    if (chain.cell_order) {
        // the scene's new positions in cell order (the runs and grid of the last substep
        // index cell slots), then the walk; hz = 0 (no grid): every lookup goes through the
        // key's cell-ordered run
        if (c.p_end > c.p_begin) {
            const uint32_t gblocks = (c.p_end - c.p_begin + kBlock - 1u) / kBlock;
            hipLaunchKernelGGL(k_gather_cell_positions, dim3(gblocks), dim3(kBlock), 0, stream,
                               (const float4 *)b.pos, b.hslot, b.pos_sorted, c.p_begin, c.p_end);
        }
        launch(GridRuns{b.grid, c.grid_base, (uint32_t)c.gx + 2u, (uint32_t)c.gy + 2u,
                        c.grid_cells && grid_current ? (uint32_t)c.gz + 2u : 0u, c.table_base, c.mask, b.crun},
               (const float4 *)b.pos_sorted);
    } else {
        launch(TableRuns{b.table, c.table_base, c.mask}, (const float4 *)b.pos);
    }
    return hipGetLastError();
}

}  // namespace sph
