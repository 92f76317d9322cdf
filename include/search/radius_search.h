// radius_search.h -- brute-force radius search (src/search/radius_search.h:11-20), on the GPU.
#pragma once

#include <cstddef>

#include "core/point_types.h"

namespace rvpoint {

// First max_nn points within radius in index order; returns min(count, max_nn).
std::size_t radius_search_sc(const PointXYZ *cloud, std::size_t n, PointXYZ query, float radius, int *indices,
                             float *dists, int max_nn);

std::size_t radius_search_rvv(const PointCloudSoA &cloud, PointXYZ query, float radius, int *indices,
                              float *dists, int max_nn);

}  // namespace rvpoint
