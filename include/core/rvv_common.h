// rvv_common.h -- the reference's distance helpers (src/core/rvv_common.h:19-32), evaluated on the GPU.
#pragma once

#include <cstddef>
#include <vector>

#include "core/point_types.h"

namespace rvpoint {

// out_d2[i] = ((x[i]-qx)^2 + (y[i]-qy)^2) + (z[i]-qz)^2 for i < n (D1, rvv_common.cpp:36-43).
void get_dist_sq_rvv(const float *x, const float *y, const float *z, float qx, float qy, float qz, float *out_d2,
                     std::size_t n);

// Appends (subset_indices[i], d2) for every subset point with d2 <= r2, in subset order
// (rvv_common.cpp:55-68).  Does not clear the outputs.
void get_inds_in_radius_rvv(const float *x, const float *y, const float *z, const int *subset_indices,
                            std::size_t n, float qx, float qy, float qz, float r2, std::vector<int> &out_indices,
                            std::vector<float> &out_dists);

}  // namespace rvpoint
