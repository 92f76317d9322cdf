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
  float icell;                 // fl(1 / cell)
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

// Cell-unit coordinate of v: fl(fl(v - mn) * icell), icell = fl(1 / cell).  Within 3 * 2^-24 (relative) of the
// exact (v - mn) / cell.
__device__ __forceinline__ float cell_coord(float v, float mn, float icell) {
  return __fmul_rn(__fsub_rn(v, mn), icell);
}

// floor(f) clamped to [0, g-1] (NaN -> 0)
__device__ __forceinline__ int floor_clamped(float f, int g) {
  if (!(f >= 0.0f)) return 0;
  if (f >= (float)(g - 1)) return g - 1;
  return (int)f;
}

// Cells [lo, hi] on one axis that hold every point p within R (1 + 3 * 2^-24) of q, the farthest a point with
// fl(d2(p, q)) <= fl(R * R) can lie.  u = cell_coord(q), rc = fl(R * icell).  The grid puts p in floor(D),
// D = fl(fl(p - mn) / cell); the roundings in D, u and rc (relative errors of at most 8 * 2^-24 together) keep
// D within (|u| + rc) * 2^-21 of [u - rc, u + rc], and the bounds below lie beyond that interval by at least
// (|u| + rc) * 0.875 * 2^-20 after their own rounding, so by monotonicity of floor every such p's cell is in
// [lo, hi].  The bounds are conservative: a window may hold one cell layer more than the exact one, never
// fewer (results do not depend on the window, only the candidates scanned).
__device__ __forceinline__ void window_axis(float u, float rc, int g, int &lo, int &hi) {
  const float m = __fmul_rn(__fadd_rn(fabsf(u), rc), 0x1p-20f);
  lo = floor_clamped(__fsub_rn(__fsub_rn(u, rc), m), g);
  hi = floor_clamped(__fadd_rn(__fadd_rn(u, rc), m), g);
}

__device__ __forceinline__ void window3(const GridView &g, float qx, float qy, float qz, float r, int lo[3],
                                        int hi[3]) {
  const float rc = __fmul_rn(r, g.icell);
  window_axis(cell_coord(qx, g.mnx, g.icell), rc, g.gx, lo[0], hi[0]);
  window_axis(cell_coord(qy, g.mny, g.icell), rc, g.gy, lo[1], hi[1]);
  window_axis(cell_coord(qz, g.mnz, g.icell), rc, g.gz, lo[2], hi[2]);
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
  Tmp xs, ys, zs, perm, rank, keys, cstart;
  hipStream_t stream = nullptr;

  // Build from device SoA.  The device grid uses the requested cell unless its dense table would exceed
  // max(2^24, 4n) cells (and always fewer than 2^32); then the edge doubles until it fits (results do not
  // depend on the edge).
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
