// pointer_octree.h -- the PointerOctree interface (src/search/pointer_octree.h:8-58) backed by the device
// hash grid.  radiusSearch and radiusSearchScalar both return the brute-force set in ascending
// (d2, index) order (the reference returns it in traversal order).  getRoot() returns a one-leaf tree: a
// root holding the cloud's bounding box, every index and the contiguous leaf coordinates, built on first
// use; the device grid takes the place of the node hierarchy.
#pragma once

#include <cstddef>
#include <memory>
#include <vector>

#include "core/point_types.h"

namespace rvpoint {

class SpatialHash;

struct PointerOctreeNode {
  float min_x, min_y, min_z;
  float max_x, max_y, max_z;
  PointerOctreeNode *children[8] = {nullptr};
  std::vector<int> indices;

  std::vector<float> leaf_x;
  std::vector<float> leaf_y;
  std::vector<float> leaf_z;

  bool is_leaf = true;

  PointerOctreeNode();
  ~PointerOctreeNode();
};

class PointerOctree {
 public:
  PointerOctree();
  ~PointerOctree();
  PointerOctree(const PointerOctree &) = delete;
  PointerOctree &operator=(const PointerOctree &) = delete;

  void setInputCloud(const PointCloudSoA &cloud);
  void build();

  std::size_t radiusSearch(const PointXYZ &query, float radius, std::vector<int> &indices,
                           std::vector<float> &dists) const;
  std::size_t radiusSearchScalar(const PointXYZ &query, float radius, std::vector<int> &indices,
                                 std::vector<float> &dists) const;

  void setMaxPointsPerLeaf(int n) { max_points_per_leaf_ = n; }
  void setMaxDepth(int d) { max_depth_ = d; }
  void setBuildEpsilon(float eps) { build_epsilon_ = eps; }

  const PointerOctreeNode *getRoot() const;

  // ---- extensions (not in the reference) ----
  const SpatialHash &hash() const;
  const PointCloudSoA &cloud() const { return cloud_; }

 private:
  PointCloudSoA cloud_{};
  std::unique_ptr<SpatialHash> hash_;
  mutable std::unique_ptr<PointerOctreeNode> root_;
  int max_points_per_leaf_ = 64;
  int max_depth_ = 8;
  float build_epsilon_ = 1e-4f;
};

}  // namespace rvpoint
