// batch.hpp -- segmented (many-cloud) grid shared by the batched stages.
#pragma once

#include <algorithm>

#include "query.hpp"

namespace rvp {

struct SegDesc {
  float mn[3];     // grid origin: the cloud's bbox min after the 0.01*edge margin
  float cell;      // grid edge
  float icell;     // fl(1 / cell)
  int g[3];        // cells per axis
  uint32_t toff;   // first cell in the shared table
};

struct SegArrays {
  const float *xs, *ys, *zs;
  const uint32_t *perm;
  const uint32_t *rank;  // batch position -> cell-order position
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
  v.rank = A.rank;
  v.cell_start = A.cstart + d.toff;
  v.mnx = d.mn[0];
  v.mny = d.mn[1];
  v.mnz = d.mn[2];
  v.cell = d.cell;
  v.icell = d.icell;
  v.gx = d.g[0];
  v.gy = d.g[1];
  v.gz = d.g[2];
  v.n = A.off[A.ncl];
  return v;
}

// Grid of the batched one-query-per-wave kernels: wave g of G takes queries g, g + G, ... of the true total
// (read from the offsets on the device), so the launch need not scale with the bound; 32768 waves are several
// rounds of the resident ones, which lets the dispatcher even out queries of uneven cost.
inline unsigned query_blocks(size_t n_bound) {
  return (unsigned)std::min<size_t>(8192, std::max<size_t>(1, ceil_div(n_bound, QW)));
}

// Grid and mode of a batch-solving query kernel (WaveBatches): strided over two rounds of resident waves
// while the batches would fill fewer than 32 rounds, else one batch of NB per wave.
template <int NB, class K> int batch_blocks(K kernel, size_t n_bound, unsigned *blocks, bool *strided) {
  unsigned res = 1;
  RVP_TRY(resident_blocks(reinterpret_cast<const void *>(kernel), QT, &res));
  const size_t batches = ceil_div(n_bound, (size_t)NB);
  *strided = batches <= (size_t)res * QW * 32;
  *blocks = *strided ? (unsigned)std::max<size_t>(1, std::min<size_t>((size_t)res * 2, ceil_div(n_bound, QW)))
                     : (unsigned)ceil_div(batches, QW);
  return RVP_OK;
}

// this wave's index g and the wave count G of the launch
__device__ __forceinline__ uint32_t wave_id() { return blockIdx.x * QW + (threadIdx.x >> 6); }
__device__ __forceinline__ uint32_t wave_count() { return gridDim.x * QW; }

// Batch k (up to NB queries) of a wave that solves a batch at a time; false when the wave is done.
// STRIDED (a grid of two rounds of resident waves, for counts that fill only a few rounds of batches): batch
// k holds the wave's queries g + (k*NB + j)*G, so every wave gets the same count to within one and no
// partial last round of whole batches is left.  Otherwise (a grid of one batch per wave, which the
// dispatcher balances): the one batch of NB neighbouring queries NB*g + j.
template <bool STRIDED, int NB>
__device__ __forceinline__ bool wave_batch(uint32_t k, uint32_t total, uint32_t &first, uint32_t &step,
                                           int &nbatch) {
  if (STRIDED) {
    const uint64_t G = wave_count(), f = (uint64_t)wave_id() + (uint64_t)k * NB * G;
    if (f >= total) return false;
    first = (uint32_t)f;
    step = (uint32_t)G;
    nbatch = (int)min<uint64_t>(NB, ((uint64_t)total - f + G - 1) / G);
    return true;
  }
  const uint64_t f = (uint64_t)wave_id() * NB;
  if (k > 0 || f >= total) return false;
  first = (uint32_t)f;
  step = 1;
  nbatch = (int)min<uint64_t>(NB, (uint64_t)total - f);
  return true;
}

// out[c] = running max of min(in[c], n_bound), out[0] = 0: the caller's offsets as a valid partition
int seg_clamp_offsets(const uint32_t *in, int ncl, size_t n_bound, uint32_t *out, hipStream_t s);

// Host owner of a segmented grid (stream-ordered temporaries).
struct SegGrid {
  Tmp desc, mm, cellv, ncells, toff, xs, ys, zs, perm, rank, keys, cstart, offs;
  int ncl = 0;
  const uint32_t *off = nullptr;  // the caller's offsets made non-decreasing, off[0] = 0, off[ncl] <= n_bound
  size_t n_bound = 0;

  SegArrays arrays() const {
    return SegArrays{xs.as<float>(), ys.as<float>(), zs.as<float>(), perm.as<uint32_t>(), rank.as<uint32_t>(),
                     cstart.as<uint32_t>(), desc.as<SegDesc>(), off, ncl};
  }

  // Build the grids of the ncl clouds [off[c], off[c+1]) of x/y/z (device offsets, n_bound >= off[ncl]).
  // The offsets are first clamped into a valid partition of [0, min(off[ncl], n_bound)); every later
  // kernel of the batch reads that copy (member off), so no index leaves the n_bound arrays.
  // mode_knn: edge from the k-NN estimate for K neighbours incl. self; otherwise edge = cell_dev[c].
  int build(const float *x, const float *y, const float *z, const uint32_t *off, int ncl, size_t n_bound,
            bool mode_knn, int K, const float *cell_dev, hipStream_t s);
};

}  // namespace rvp
