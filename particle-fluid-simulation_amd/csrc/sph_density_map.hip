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

template <class Runs>
__global__ __launch_bounds__(kBlock) void k_density_map(SceneConst c, const float4 *__restrict__ pos,
                                                       Runs runs, MapGeometry g,
                                                       float *__restrict__ out) {
    const uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= g.voxels) return;
    const int vx = (int)(v % g.dx), vy = (int)((v / g.dx) % g.dy), vz = (int)(v / ((uint64_t)g.dx * g.dy));
    if (map_out_of_bounds(vx, vy, vz, g)) {
        out[v] = 0.0f;
        return;
    }
    const float cx = c.minx + ((float)vx + 0.5f) * g.voxel;
    const float cy = c.miny + ((float)vy + 0.5f) * g.voxel;
    const float cz = c.minz + ((float)vz + 0.5f) * g.voxel;
    // samplePointToCellCoord: -1 on every axis when the centre lies outside the box
    const bool outside = cx < c.minx || cy < c.miny || cz < c.minz || cx > c.maxx || cy > c.maxy ||
                         cz > c.maxz;
    const int ox = outside ? -1 : cell_of(cx, c.minx, c.R);
    const int oy = outside ? -1 : cell_of(cy, c.miny, c.R);
    const int oz = outside ? -1 : cell_of(cz, c.minz, c.R);
    const float R = c.R, R2 = c.R2;
    float density = 0.0f;
    constexpr int kU = 4;  // candidates of a run loaded together
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = oy + dy, z = oz + dz;
            if (y < 0 || z < 0 || y > g.my || z > g.mz) continue;  // the whole row is out
            int2 se[3];
            runs.row3(ox, y, z, se);  // the row's three cells, loaded together
            for (int dx = 0; dx < 3; ++dx) {
                if (map_out_of_bounds(ox + dx - 1, y, z, g)) continue;
                if (se[dx].x == -1) continue;
                const float4 *src = pos + se[dx].x;
                const int cnt = se[dx].y - se[dx].x;
                for (int k = 0; k < cnt; k += kU) {
                    float4 p[kU];
                    load_candidates<kU>(src, k, cnt, p);
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        if (k + u >= cnt) continue;
                        const float ex = cx - p[u].x, ey = cy - p[u].y, ez = cz - p[u].z;
                        const float d2 = ex * ex + ey * ey + ez * ez;
                        if (d2 < R2) {
                            const float dist = sqrtf(d2);
                            const float t = dist >= R ? 0.0f : (R - dist) * (R - dist);
                            density = density + c.m * t;
                        }
                    }
                }
            }
        }
    density = density * 1.0f;
    out[v] = density / c.volume;
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
    const uint64_t blocks = (g.voxels + kBlock - 1u) / kBlock;
    const ChainBuffers &b = chain.b;
    if (chain.cell_order) {
        // hz = 0 (no grid): every lookup goes through the key's cell-ordered run
        const GridRuns runs{b.grid, c.grid_base, (uint32_t)c.gx + 2u, (uint32_t)c.gy + 2u,
                            c.grid_cells && grid_current ? (uint32_t)c.gz + 2u : 0u, c.table_base, c.mask,
                            b.crun};
        hipLaunchKernelGGL(k_density_map<GridRuns>, dim3((uint32_t)blocks), dim3(kBlock), 0, stream, c,
                           (const float4 *)b.pos_sorted, runs, g, out);
    } else {
        const TableRuns runs{b.table, c.table_base, c.mask};
        hipLaunchKernelGGL(k_density_map<TableRuns>, dim3((uint32_t)blocks), dim3(kBlock), 0, stream, c,
                           (const float4 *)b.pos, runs, g, out);
    }
    return hipGetLastError();
}

}  // namespace sph
