// point_types.h -- point layouts of the drop-in API (same layout and names as the reference's
// src/core/point_types.h:12-29, so existing callers compile unchanged).
#pragma once

#include <cstddef>
#include <vector>

namespace rvpoint {

// One point, 12 bytes (x, y, z), as the reference's PointXYZ.
struct PointXYZ {
  float x, y, z;
};

// Non-owning structure-of-arrays view: the caller owns x/y/z and keeps them alive while a SpatialHash
// built on the view exists.
struct PointCloudSoA {
  float *x = nullptr;
  float *y = nullptr;
  float *z = nullptr;
  std::size_t n = 0;
};

}  // namespace rvpoint
