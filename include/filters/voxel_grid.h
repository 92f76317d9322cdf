// voxel_grid.h -- voxel-grid downsampling (src/filters/voxel_grid.h:14-34), on the GPU.
#pragma once

#include <cstddef>

#include "core/point_types.h"

namespace rvpoint {

// Centroid of every occupied voxel floor(p / leaf), in lexicographic (vx, vy, vz) order; returns the count.
std::size_t voxel_grid_downsamp_sc(const PointXYZ *in, std::size_t n, PointXYZ *out, float leaf_size);

// Same result as voxel_grid_downsamp_sc (the reference's x86 build routes here too).
std::size_t voxel_grid_downsamp_rvv(const PointCloudSoA &in, PointXYZ *out, float leaf_size);

std::size_t voxel_grid_downsamp_rvv_v2(const PointCloudSoA &in, PointXYZ *out, float leaf_size);

}  // namespace rvpoint
