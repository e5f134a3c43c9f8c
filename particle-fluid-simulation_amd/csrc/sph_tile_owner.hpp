// Cell tiles of the tiled neighbour passes: which tile owns a particle, and the geometry
// of a tile's staged region.
//
// A scene's box grid is covered by tiles of tw x th x td cells.  A tile's region is the
// tile grown by one cell per side: it holds every cell of the 27-cell neighbourhood of any
// cell inside the tile.  A particle belongs to the tile of its predicted cell c0 (the cell
// both neighbour passes walk around) when
//   * its predicted position lies inside the box (c0 is then inside the grid),
//   * the density pass's sample cells are exactly c0 - 1, c0, c0 + 1 on every axis where
//     the sample point is inside the box (calculateDensities.comp:114-118), so that one
//     walk over the force pass's 27 cells serves both passes, and
//   * its current cell (whose key run holds it, updateCellEntries.comp:73-77) lies within
//     one cell of c0, so that the run is staged in the tile's region.
// The region cell of the current cell is the particle's home copy: a run staged for two
// region cells whose keys collide holds the particle twice, and only the copy at its own
// current cell selects it.  Every other particle is a stray of the tiled launches.
#pragma once

#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

struct TileRegion {
    int32_t ox, oy, oz;  // first region cell: the tile's first cell minus one per axis
    uint32_t rx, ry, rz; // region extent in cells (tile + 2 per axis)
    uint32_t cells;      // rx * ry * rz
};

__device__ __forceinline__ TileRegion tile_region(const SceneConst &c, uint32_t t_local) {
    const uint32_t tx = t_local % c.ntx;
    const uint32_t rest = t_local / c.ntx;
    const uint32_t ty = rest % c.nty, tz = rest / c.nty;
    TileRegion r;
    r.ox = (int32_t)(tx * c.tw) - 1;
    r.oy = (int32_t)(ty * c.th) - 1;
    r.oz = (int32_t)(tz * c.td) - 1;
    r.rx = c.tw + 2u;
    r.ry = c.th + 2u;
    r.rz = c.td + 2u;
    r.cells = r.rx * r.ry * r.rz;
    return r;
}

// (owning tile, region cell of the home copy) of a particle at current position P with
// predicted position s, or (kNoTile, 0).
__device__ __forceinline__ uint2 tile_owner(const SceneConst &c, const float3 &P, const float3 &s) {
    if (!c.use_tiles) return make_uint2(kNoTile, 0u);
    const AxisCells ax = sample_axis(s.x, c.minx, c.maxx, c.R);
    const AxisCells ay = sample_axis(s.y, c.miny, c.maxy, c.R);
    const AxisCells az = sample_axis(s.z, c.minz, c.maxz, c.R);
    auto same = [](const AxisCells &a) {
        return a.in(1) && (!a.in(0) || a.lo_off == 0xFFFFFFFFu) && (!a.in(2) || a.hi_off == 1u);
    };
    if (!(same(ax) && same(ay) && same(az))) return make_uint2(kNoTile, 0u);
    // in(1) on every axis: s is inside the box, so c0 lies in [0, g) per axis
    const int32_t x0 = (int32_t)ax.mid, y0 = (int32_t)ay.mid, z0 = (int32_t)az.mid;
    const int32_t dx = cell_of(P.x, c.minx, c.R) - x0;
    const int32_t dy = cell_of(P.y, c.miny, c.R) - y0;
    const int32_t dz = cell_of(P.z, c.minz, c.R) - z0;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return make_uint2(kNoTile, 0u);
    const uint32_t tx = (uint32_t)x0 / c.tw, ty = (uint32_t)y0 / c.th, tz = (uint32_t)z0 / c.td;
    const uint32_t t = c.tile_base + tx + c.ntx * (ty + c.nty * tz);
    // region coordinates of the current cell: (x0 + dx) - (tx * tw - 1)
    const uint32_t lx = (uint32_t)(x0 - (int32_t)(tx * c.tw) + 1 + dx);
    const uint32_t ly = (uint32_t)(y0 - (int32_t)(ty * c.th) + 1 + dy);
    const uint32_t lz = (uint32_t)(z0 - (int32_t)(tz * c.td) + 1 + dz);
    return make_uint2(t, lx + (c.tw + 2u) * (ly + (c.th + 2u) * lz));
}

}  // namespace sph
