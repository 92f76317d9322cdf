// grid.hpp -- the device-resident uniform grid behind SpatialHash (cell-sorted SoA + dense cell table).
#pragma once

#include <cfloat>

#include "common.hpp"

namespace rvp {

// Kernel-side view of a built grid (passed by value).
struct GridView {
  const float *xs, *ys, *zs;   // points in cell order (SoA)
  const uint32_t *perm;        // cell-order position -> original index
  const uint32_t *rank;        // original index -> cell-order position (the inverse of perm)
  const uint32_t *cell_start;  // [ncells + 1]: points of cell c are [cell_start[c], cell_start[c+1])
  float mnx, mny, mnz;         // grid origin = bbox min after the 0.01*cell margin (spatial_hashing.cpp:141-143)
  float cell;                  // cell edge of this grid (>= the requested one, see Grid::build)
  int gx, gy, gz;              // cells per axis; the linear cell id is ix + gx*(iy + gy*iz) (x fastest)
  uint32_t n;
};

// Cell coordinate of v on one axis: floor((v - mn) / cell) with IEEE division (spatial_hashing.h:49-51),
// clamped to [0, g-1].  Monotone non-decreasing in v, which is all the queries rely on.
__device__ __forceinline__ int cell_clamped(float v, float mn, float c, int g) {
  float f = floorf(__fdiv_rn(__fsub_rn(v, mn), c));
  if (!(f >= 0.0f)) return 0;
  if (f >= (float)(g - 1)) return g - 1;
  return (int)f;
}

// Cells [lo, hi] on one axis that can hold a point p with fl(d2(p, q)) <= fl(r*r).
// Such a point has |p - q| <= r (1 + 3u) (u = 2^-24); the bounds below are floats strictly beyond
// q -/+ r (1 + 2^-20), so by monotonicity of cell_clamped every such p lies in [lo, hi].
__device__ __forceinline__ void window_axis(float q, float r, float mn, float c, int g, int &lo, int &hi) {
  const double rr = (double)r * (1.0 + 0x1p-20);
  float bl = (float)((double)q - rr);
  float bh = (float)((double)q + rr);
  bl = next_down(bl);
  bh = next_up(bh);
  lo = cell_clamped(bl, mn, c, g);
  hi = cell_clamped(bh, mn, c, g);
}

__device__ __forceinline__ void window3(const GridView &g, float qx, float qy, float qz, float r, int lo[3],
                                        int hi[3]) {
  window_axis(qx, r, g.mnx, g.cell, g.gx, lo[0], hi[0]);
  window_axis(qy, r, g.mny, g.cell, g.gy, lo[1], hi[1]);
  window_axis(qz, r, g.mnz, g.cell, g.gz, lo[2], hi[2]);
}

// Host-side owner of a grid's device buffers (stream-ordered pool memory).
struct Grid {
  // reference-visible state (SpatialHash::setInputCloud)
  const float *x = nullptr, *y = nullptr, *z = nullptr;
  size_t n = 0;
  float req_cell = 0.f;     // cell size passed by the caller
  float bb[6] = {0, 0, 0, 0, 0, 0};
  // device grid
  float cell = 0.f;         // cell size of the device grid
  int dims[3] = {0, 0, 0};
  uint64_t ncells = 0;
  bool built = false;
  bool persistent = false;  // owned by an rvp_hash: hipMalloc'd buffers that outlive the build stream
  Tmp xs, ys, zs, perm, rank, keys, cstart, occ;
  hipStream_t stream = nullptr;

  // Build from device SoA.  The device grid uses the requested cell unless its dense table would exceed
  // max(2^24, 4n) cells; then the edge doubles until it fits (results do not depend on the edge).
  int build(const float *x, const float *y, const float *z, size_t n, float cell, hipStream_t s);
  // Build with a cell edge chosen for k-nearest-neighbour queries.
  int build_for_knn(const float *x, const float *y, const float *z, size_t n, int k, hipStream_t s);
  // Build from the data min/max (device bbox already read back): one host sync fewer.
  int build_from(const float mm[6], float cell, hipStream_t s);
  int minmax_host(float mm[6], hipStream_t s);
  size_t occupied_cells() const;  // synchronises
  GridView view() const;
};

// device-wide bbox (min/max ignoring NaN) -> {min[3], max[3]} of the data (no margin), device out
int bbox_minmax(const float *x, const float *y, const float *z, size_t n, float *out6_dev, hipStream_t s);

}  // namespace rvp

struct rvp_hash {
  rvp::Grid g;
};
