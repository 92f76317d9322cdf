// statistical_outlier_removal.h -- SOR (src/filters/statistical_outlier_removal.h:18-42), on the GPU.
#pragma once

#include <cstddef>

#include "core/point_types.h"

namespace rvpoint {

class SpatialHash;
class Octree;
class PointerOctree;

// Exact k-NN statistical outlier removal; kept points in input order; returns the count.
std::size_t sor_sc(const PointXYZ *in, std::size_t n, PointXYZ *out, int k, float alpha);
std::size_t sor_rvv(const PointCloudSoA &in, PointXYZ *out, int k, float alpha);

// Radius-bounded variant over a SpatialHash built on the same cloud.
std::size_t sor_spatial_hash(const PointCloudSoA &in, const SpatialHash &hash, PointXYZ *out, int k, float alpha,
                             float search_radius = 0.5f);

// The same over the tree facades: all neighbours within the radius, the k smallest after the first, summed
// ascending (the reference's PointerOctree variant sums them in nth_element's unspecified order).
std::size_t sor_octree(const PointCloudSoA &in, const Octree &tree, PointXYZ *out, int k, float alpha,
                       float search_radius = 0.5f);
std::size_t sor_pointer_octree(const PointCloudSoA &in, const PointerOctree &tree, PointXYZ *out, int k,
                               float alpha, float search_radius = 0.5f);

}  // namespace rvpoint
