// Density map (updateDensityMap.comp) launcher shared by the world and batch entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "sph_device.hpp"
#include "sph_host.hpp"

namespace sph {

// Voxel grid of the density map over one scene's box.
struct MapGeometry {
    float voxel;             // voxel edge (DENSITY_CELL_SIZE)
    uint32_t dx, dy, dz;     // voxels per axis: the image size (renderer.c:774-783)
    int32_t mx, my, mz;      // the shader's bound, ceil((max - min) / voxel) (:95)
    uint64_t voxels;         // dx * dy * dz
};

// Fills g for the scene's box; SPH_ERR_INVALID_ARG / SPH_ERR_CAPACITY on a bad voxel size.
int map_geometry(const sph_params &p, float voxel_size, MapGeometry &g);

// Writes g.voxels floats (x fastest) to device memory `out` on `stream`, from the state
// and cell table `chain` holds after at least one substep of the scene `c`.  grid_current:
// the halo grid was filled for c's cell grid (else runs are found through their keys).
hipError_t launch_density_map(hipStream_t stream, const SceneConst &c, const Chain &chain,
                              bool grid_current, const MapGeometry &g, float *out);

}  // namespace sph
