// ransac_plane.h -- RANSAC plane fit and plane split (src/segmentation/ransac_plane.h:11-45), on the GPU.
// Like the reference, ransac_plane_* reseed and advance the C library's rand() (srand(0), three rand()
// calls per iteration).
#pragma once

#include <cstddef>

#include "core/point_types.h"

namespace rvpoint {

int ransac_plane_sc(const PointXYZ *cloud, std::size_t n, float dist_thresh, int max_iters, float *model,
                    float collinear_thresh = 1e-6f, float probability = 0.99f);

int ransac_plane_rvv(const PointCloudSoA &cloud, float dist_thresh, int max_iters, float *model,
                     float collinear_thresh = 1e-6f, float probability = 0.99f);

std::size_t extract_plane_inliers_rvv(const PointCloudSoA &cloud, const float *model, float dist_thresh,
                                      PointXYZ *inliers);

std::size_t extract_plane_outliers_rvv(const PointCloudSoA &cloud, const float *model, float dist_thresh,
                                       PointXYZ *outliers);

void extract_plane_inliers_outliers_rvv(const PointCloudSoA &cloud, const float *model, float dist_thresh,
                                        PointXYZ *inliers, PointXYZ *outliers, std::size_t &n_inliers,
                                        std::size_t &n_outliers);

}  // namespace rvpoint
