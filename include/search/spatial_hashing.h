// spatial_hashing.h -- SpatialHash with the reference's interface (src/search/spatial_hashing.h:12-53);
// the grid lives in HBM and queries run on the GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "core/point_types.h"

struct rvp_hash;

namespace rvpoint {

class SpatialHash {
 public:
  SpatialHash();
  ~SpatialHash();
  SpatialHash(const SpatialHash &) = delete;
  SpatialHash &operator=(const SpatialHash &) = delete;

  // Keeps the view; build() copies the cloud to the device and builds the grid.
  void setInputCloud(const PointCloudSoA &cloud, float cell_size);
  void build();

  // Neighbours within radius, ascending (d2, index); truncated to max_nn when max_nn > 0.
  std::size_t radiusSearch(const PointXYZ &query, float radius, std::vector<int> &indices,
                           std::vector<float> &dists, int max_nn = 0) const;

  // Accepted for compatibility: the device grid does not hash cells, so results never depend on them.
  void setHashPrimes(int64_t p1, int64_t p2) {
    p1_ = p1;
    p2_ = p2;
  }

  // ---- batched extensions (not in the reference) ----
  // All points' lists in one launch: rows of stride max_nn (1..64 fast path); counts per point.
  void radiusSearchAll(float radius, int max_nn, std::vector<int> &indices, std::vector<float> &dists,
                       std::vector<int> &counts) const;
  const PointCloudSoA &cloud() const { return cloud_; }
  rvp_hash *handle() const { return h_; }
  const float *deviceX() const { return dx_; }
  const float *deviceY() const { return dy_; }
  const float *deviceZ() const { return dz_; }

 private:
  PointCloudSoA cloud_;
  float cell_size_ = 0.0f;
  int64_t p1_ = 73856093LL;
  int64_t p2_ = 19349663LL;
  rvp_hash *h_ = nullptr;
  float *dx_ = nullptr, *dy_ = nullptr, *dz_ = nullptr;  // device copy of the cloud
  std::size_t dev_n_ = 0;
};

}  // namespace rvpoint
