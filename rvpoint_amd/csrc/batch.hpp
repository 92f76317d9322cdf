// batch.hpp -- segmented (many-cloud) grid shared by the batched stages.
#pragma once

#include <algorithm>

#include "query.hpp"

namespace rvp {

struct SegDesc {
  float mn[3];     // grid origin: the cloud's bbox min after the 0.01*edge margin
  float cell;      // grid edge
  int g[3];        // cells per axis
  uint32_t toff;   // first cell in the shared table
};

struct SegArrays {
  const float *xs, *ys, *zs;
  const uint32_t *perm;
  const uint32_t *cstart;
  const SegDesc *desc;
  const uint32_t *off;  // [ncl + 1]
  int ncl;
};

__device__ __forceinline__ int cloud_of(const uint32_t *__restrict__ off, int ncl, uint32_t i) {
  int lo = 0, hi = ncl - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ GridView seg_view(const SegArrays &A, int c) {
  const SegDesc d = A.desc[c];
  GridView v;
  v.xs = A.xs;
  v.ys = A.ys;
  v.zs = A.zs;
  v.perm = A.perm;
  v.cell_start = A.cstart + d.toff;
  v.mnx = d.mn[0];
  v.mny = d.mn[1];
  v.mnz = d.mn[2];
  v.cell = d.cell;
  v.gx = d.g[0];
  v.gy = d.g[1];
  v.gz = d.g[2];
  v.n = A.off[A.ncl];
  return v;
}

// grid of the batched query kernels: they stride over the true total read from the offsets, so the
// launch need not scale with the bound
inline unsigned query_blocks(size_t n_bound, size_t per_wave = 1) {
  return (unsigned)std::min<size_t>(8192, std::max<size_t>(1, ceil_div(n_bound, QW * per_wave)));
}

// out[c] = running max of min(in[c], n_bound), out[0] = 0: the caller's offsets as a valid partition
int seg_clamp_offsets(const uint32_t *in, int ncl, size_t n_bound, uint32_t *out, hipStream_t s);

// Host owner of a segmented grid (stream-ordered temporaries).
struct SegGrid {
  Tmp desc, mm, cellv, ncells, toff, xs, ys, zs, perm, keys, cstart, offs;
  int ncl = 0;
  const uint32_t *off = nullptr;  // the caller's offsets made non-decreasing, off[0] = 0, off[ncl] <= n_bound
  size_t n_bound = 0;

  SegArrays arrays() const {
    return SegArrays{xs.as<float>(), ys.as<float>(), zs.as<float>(), perm.as<uint32_t>(), cstart.as<uint32_t>(),
                     desc.as<SegDesc>(), off, ncl};
  }

  // Build the grids of the ncl clouds [off[c], off[c+1]) of x/y/z (device offsets, n_bound >= off[ncl]).
  // The offsets are first clamped into a valid partition of [0, min(off[ncl], n_bound)); every later
  // kernel of the batch reads that copy (member off), so no index leaves the n_bound arrays.
  // mode_knn: edge from the k-NN estimate for K neighbours incl. self; otherwise edge = cell_dev[c].
  int build(const float *x, const float *y, const float *z, const uint32_t *off, int ncl, size_t n_bound,
            bool mode_knn, int K, const float *cell_dev, hipStream_t s);
};

}  // namespace rvp
