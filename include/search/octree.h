// octree.h -- the Octree interface (src/search/octree.h:11-48) backed by the device hash grid.
//
// radiusSearch returns exactly the brute-force set {j : D1(j) <= fl(r*r)}, as the reference's tree does;
// the list comes in ascending (d2, index) order instead of the tree's traversal order, which depends on
// how the tree happened to split.  max_nn is ignored, as in the reference (octree.cpp:111-120).  The
// leaf-size, depth and epsilon setters are kept for source compatibility: the device grid picks its own
// cell edge, and results never depend on it.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "core/point_types.h"

namespace rvpoint {

class SpatialHash;

struct OctreeNode {
  float min_x, min_y, min_z;
  float max_x, max_y, max_z;
  OctreeNode *children[8] = {nullptr};
  std::vector<int> indices;
  bool is_leaf = true;

  ~OctreeNode();
};

class Octree {
 public:
  Octree();
  ~Octree();
  Octree(const Octree &) = delete;
  Octree &operator=(const Octree &) = delete;

  void setInputCloud(const PointCloudSoA &cloud);
  void build();

  std::size_t radiusSearch(const PointXYZ &query, float radius, std::vector<int> &indices,
                           std::vector<float> &dists, int max_nn = 0) const;

  void setMaxPointsPerLeaf(int n) { max_points_per_leaf_ = n; }
  void setMaxDepth(int d) { max_depth_ = d; }
  void setBuildEpsilon(float eps) { build_epsilon_ = eps; }

  // ---- extensions (not in the reference) ----
  const SpatialHash &hash() const;  // the device grid the searches run on (valid after build())
  const PointCloudSoA &cloud() const { return cloud_; }

 private:
  PointCloudSoA cloud_{};
  std::unique_ptr<SpatialHash> hash_;
  int max_points_per_leaf_ = 64;
  int max_depth_ = 8;
  float build_epsilon_ = 1e-4f;
};

}  // namespace rvpoint
