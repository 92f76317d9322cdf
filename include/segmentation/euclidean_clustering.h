// euclidean_clustering.h -- Euclidean cluster extraction (src/segmentation/euclidean_clustering.h:16-65) on
// the GPU: connected components of the radius graph, listed by smallest index, each ascending -- the
// clusters the reference's breadth-first search grows.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <vector>

#include "core/point_types.h"

namespace rvpoint {

class Octree;
class PointerOctree;

struct ClusterIndices {
  std::vector<int> indices;  // point indices into the source cloud
};

class EuclideanClustering {
 public:
  EuclideanClustering() = default;

  void setInputCloud(const PointCloudSoA &cloud) { cloud_ = &cloud; }
  void setClusterTolerance(float tolerance) { clusterTolerance_ = tolerance; }
  void setMinClusterSize(int min_size) { minClusterSize_ = min_size; }
  void setMaxClusterSize(int max_size) { maxClusterSize_ = max_size; }

  // Accepted for compatibility: the GPU grid answers the radius queries, with the same neighbour sets as the
  // reference's search structures.
  void setNeighborSearch(const Octree *) {}
  void setNeighborSearch(const PointerOctree *) {}

  float clusterTolerance() const { return clusterTolerance_; }
  int minClusterSize() const { return minClusterSize_; }
  int maxClusterSize() const { return maxClusterSize_; }

  std::vector<ClusterIndices> extract() const;

 private:
  const PointCloudSoA *cloud_ = nullptr;
  float clusterTolerance_ = 0.05f;
  int minClusterSize_ = 1;
  int maxClusterSize_ = std::numeric_limits<int>::max();
};

}  // namespace rvpoint
