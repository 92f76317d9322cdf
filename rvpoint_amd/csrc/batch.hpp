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
__device__ __forceinline__ uint32_t wave_id() { return blockIdx.x * QW + (uint32_t)wave_in_block(); }
__device__ __forceinline__ uint32_t wave_count() { return gridDim.x * QW; }

// Work of the query kernels kept on one XCD.  The units [0, units) are split into NXCD contiguous ranges, one
// per XCD (workgroups are dispatched round-robin over the XCDs, so block b runs on XCD b mod NXCD); neighbouring
// units (queries in cell order) then stay on one XCD, whose L2 holds only its share of the cloud.
constexpr int NXCD = 8;
constexpr int XCD_PAD = 32;  // 32-bit words between the counters: one 128-byte line each, no shared line

__device__ __forceinline__ uint32_t xcd_lo(uint32_t units, int x) {
  return (uint32_t)((uint64_t)units * (uint64_t)x / NXCD);
}

// Mode 2: without atomics.  Wave i of XCD x's waves takes units lo_x + i, lo_x + i + G_x, ... of that range
// (G_x = the launch's waves on XCD x).
struct XcdStride {
  uint32_t u, hi, step;
  __device__ __forceinline__ void init(uint32_t units) {
    const int x = (int)(blockIdx.x % NXCD);
    const uint32_t blocks_x = (gridDim.x - (uint32_t)x + NXCD - 1) / NXCD;  // blocks of this launch on XCD x
    const uint32_t wpb = blockDim.x / WAVE;
    u = xcd_lo(units, x) + (blockIdx.x / NXCD) * wpb + (uint32_t)wave_in_block();
    hi = xcd_lo(units, x + 1);
    step = blocks_x * wpb;
  }
  __device__ __forceinline__ bool next(uint32_t &unit) {
    if (u >= hi) return false;
    unit = u;
    u += step;
    return true;
  }
};

// Mode 1: persistent waves that steal.  A wave starts on its own XCD's range, takes `chunk` units at a time
// from that range's counter, and when the range is drained moves on to the next XCD's range, until every range
// is drained.  q: the stream's work counters (NXCD take counters XCD_PAD words apart, then the count of
// finished waves); the last wave of the launch to finish clears them, so every launch starts from zero (also
// when replayed from a captured graph).
struct XcdWork {
  uint32_t *q;
  uint32_t units, chunk;
  int home, probe;
  __device__ __forceinline__ void init(uint32_t *q_, uint32_t units_, uint32_t chunk_) {
    q = q_;
    units = units_;
    chunk = chunk_;
    home = (int)(blockIdx.x % NXCD);
    probe = 0;
  }
  // the wave's next units [u0, u1) (wave-uniform); false once every range is drained
  __device__ __forceinline__ bool next(uint32_t &u0, uint32_t &u1) {
    while (probe < NXCD) {
      const int x = (home + probe) % NXCD;
      const uint32_t lo = xcd_lo(units, x), hi = xcd_lo(units, x + 1);
      uint32_t t = 0;
      if (lane() == 0) t = atomicAdd(&q[x * XCD_PAD], chunk);
      t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
      const uint64_t start = (uint64_t)lo + t;
      if (start < hi) {
        u0 = (uint32_t)start;
        u1 = (uint32_t)min<uint64_t>(start + chunk, hi);
        return true;
      }
      ++probe;
    }
    return false;
  }
  // every wave of the launch calls this once, after its last next(); the last one clears the counters (all other
  // waves have finished their takes by then)
  __device__ __forceinline__ void finish() {
    if (lane() == 0) {
      const uint32_t waves = gridDim.x * (blockDim.x / WAVE);
      if (atomicAdd(&q[NXCD * XCD_PAD], 1u) == waves - 1) {
        for (int x = 0; x <= NXCD; ++x)
          __hip_atomic_store(&q[x * XCD_PAD], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
};

// Batch k (up to NB queries) of a wave that solves a batch at a time; false when the wave is done.
// STRIDED (a grid of two rounds of resident waves, for counts that fill only a few rounds of batches): batch
// k holds the wave's queries g + (k*nb + j)*G, so every wave gets the same count to within one and no
// partial last round of whole batches is left.  Otherwise (a grid of one batch per wave, which the
// dispatcher balances): the one batch of nb neighbouring queries nb*g + j.
template <bool STRIDED>
__device__ __forceinline__ bool wave_batch(uint32_t k, uint32_t total, int nb, uint32_t &first, uint32_t &step,
                                           int &nbatch) {
  if (STRIDED) {
    const uint64_t G = wave_count(), f = (uint64_t)wave_id() + (uint64_t)k * (uint64_t)nb * G;
    if (f >= total) return false;
    first = (uint32_t)f;
    step = (uint32_t)G;
    nbatch = (int)min<uint64_t>((uint64_t)nb, ((uint64_t)total - f + G - 1) / G);
    return true;
  }
  const uint64_t f = (uint64_t)wave_id() * (uint64_t)nb;
  if (k > 0 || f >= total) return false;
  first = (uint32_t)f;
  step = 1;
  nbatch = (int)min<uint64_t>((uint64_t)nb, (uint64_t)total - f);
  return true;
}

// out[c] = running max of min(in[c], n_bound), out[0] = 0: the caller's offsets as a valid partition
int seg_clamp_offsets(const uint32_t *in, int ncl, size_t n_bound, uint32_t *out, hipStream_t s);

// Host owner of a segmented grid (stream-ordered temporaries).
struct SegGrid {
  Tmp desc, mm, cellv, ncells, toff, xs, ys, zs, perm, rank, cstart, offs;
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
