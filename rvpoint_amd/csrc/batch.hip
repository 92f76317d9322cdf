// batch.hip -- many clouds through one launch per stage (the batched workload of the north star).
//
// A batch is one SoA array with the clouds back to back and device offsets off[ncl+1].  Each cloud gets its
// own grid (origin, edge, dims), all grids share one cell table: cloud c's cells occupy
// [toff[c], toff[c] + ncells_c), so the sort key of a point is toff[c] + its local cell id, one stable
// radix sort orders the whole batch cloud-major and cell-major, and a query builds its cloud's GridView from
// the descriptor table and runs the same wave kernels as a single cloud (query.hpp).  Results per cloud are
// identical to processing that cloud alone.  Counts stay in device memory: every kernel is launched for an
// upper bound of points and reads the true totals from the offsets, so a stage never waits on the host.
#include <algorithm>
#include <cmath>

#include "batch.hpp"
#include "prims.hpp"

namespace rvp {

// ------------------------------------------------------------------------------------------------
// segmented grid build
// ------------------------------------------------------------------------------------------------
// out[c] = running max of min(in[c], n_bound), out[0] = 0: one wave, 64 offsets per step
__global__ __launch_bounds__(64) void seg_offsets_k(const uint32_t *__restrict__ in, int ncl, uint32_t n_bound,
                                                    uint32_t *__restrict__ out) {
  const int l = lane();
  uint32_t carry = 0;
  for (int b = 0; b <= ncl; b += 64) {
    const int c = b + l;
    uint32_t v = c == 0 ? 0u : (c <= ncl ? min(in[c], n_bound) : n_bound);
    v = max(v, carry);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(v, d);
      if (l >= d) v = max(v, o);
    }
    if (c <= ncl) out[c] = v;
    carry = __shfl(v, 63);
  }
}

// float <-> uint order-preserving map (for atomic min/max of finite floats)
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// The first launch of a grid build: the caller's offsets made a valid partition (wave 0 of block 0, as
// seg_offsets_k), the integer bbox at its identity and `words` 32-bit words of `zero` cleared.
__global__ __launch_bounds__(256) void seg_prep_k(const uint32_t *__restrict__ in, int ncl, uint32_t n_bound,
                                                  uint32_t *__restrict__ off, uint32_t *__restrict__ mmu,
                                                  uint32_t *__restrict__ zero, size_t words) {
  if (blockIdx.x == 0 && threadIdx.x < WAVE) {
    const int l = lane();
    uint32_t carry = 0;
    for (int b = 0; b <= ncl; b += WAVE) {
      const int c = b + l;
      uint32_t v = c == 0 ? 0u : (c <= ncl ? min(in[c], n_bound) : n_bound);
      v = max(v, carry);
#pragma unroll
      for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t o = bperm(v, up_addr(d));
        if (l >= d) v = max(v, o);
      }
      if (c <= ncl) off[c] = v;
      carry = readlaneu(v, WAVE - 1);
    }
  }
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (size_t)6 * ncl; i += stride)
    mmu[i] = f2ord((i % 6) < 3 ? FLT_MAX : -FLT_MAX);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) zero[i] = 0u;
}

// per-cloud bbox: blockIdx.y = cloud, blockIdx.x strides over its points; strict comparisons as the
// reference's scalar bbox (NaN never updates, spatial_hashing.cpp:75-84)
__global__ __launch_bounds__(256) void seg_minmax_k(const float *__restrict__ x, const float *__restrict__ y,
                                                     const float *__restrict__ z, const uint32_t *__restrict__ off,
                                                     uint32_t *__restrict__ mmu) {
  const int c = blockIdx.y;
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  const uint32_t e = off[c + 1];
  for (uint32_t i = off[c] + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    const float a = x[i], b = y[i], cz = z[i];
    if (a < v[0]) v[0] = a;
    if (b < v[1]) v[1] = b;
    if (cz < v[2]) v[2] = cz;
    if (a > v[3]) v[3] = a;
    if (b > v[4]) v[4] = b;
    if (cz > v[5]) v[5] = cz;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1)
    for (int j = 0; j < 3; ++j) {
      const float o = shfl_xor(v[j], d);
      if (o < v[j]) v[j] = o;
      const float p = shfl_xor(v[j + 3], d);
      if (p > v[j + 3]) v[j + 3] = p;
    }
  __shared__ float sm[4][6];
  if (lane() == 0)
    for (int j = 0; j < 6; ++j) sm[threadIdx.x >> 6][j] = v[j];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    float r = sm[0][j];
    for (int w = 1; w < 4; ++w) r = j < 3 ? fminf(r, sm[w][j]) : fmaxf(r, sm[w][j]);
    if (j < 3) atomicMin(&mmu[6 * c + j], f2ord(r));
    else atomicMax(&mmu[6 * c + j], f2ord(r));
  }
}

__global__ void seg_minmax_final_k(int ncl, const uint32_t *__restrict__ mmu, float *__restrict__ mm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 6 * ncl) mm[i] = ord2f(mmu[i]);
}

int seg_clamp_offsets(const uint32_t *in, int ncl, size_t n_bound, uint32_t *out, hipStream_t s) {
  RVP_LAUNCH("seg_offsets_k", s, seg_offsets_k<<<1, 64, 0, s>>>(in, ncl, (uint32_t)n_bound, out));
  return RVP_OK;
}

// cells per axis for an edge: floor((max - origin) / edge) + 1 (the device cell function's arithmetic)
__device__ __forceinline__ double axis_cells(float mx, float origin, float edge) {
  return (double)floorf((mx - origin) / edge) + 1.0;
}

constexpr int OCC_BITS = 1 << 16;  // occupancy cells per cloud (one byte each) for the k-NN edge estimate
constexpr uint32_t SEG_MIN_CELLS = 1u << 17;  // dense-table cells every cloud may use at least

// k-NN edge, step 1: volume-based edge c0 (coarsened until its grid fits the occupancy bitmap)
__global__ void seg_knn_c0_k(const uint32_t *__restrict__ mmu, float *__restrict__ mm, const uint32_t *__restrict__ off,
                             int ncl, int K, float *__restrict__ c0, SegDesc *__restrict__ d0) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  float m[6];  // the cloud's bbox (as seg_minmax_final_k), kept in mm for the later stages
  for (int j = 0; j < 6; ++j) {
    m[j] = ord2f(mmu[6 * c + j]);
    mm[6 * c + j] = m[j];
  }
  const double n = (double)(off[c + 1] - off[c]);
  double ext[3], emax = 0.0, vol = 1.0;
  for (int j = 0; j < 3; ++j) {
    ext[j] = m[j + 3] >= m[j] ? (double)m[j + 3] - (double)m[j] : 0.0;
    emax = fmax(emax, ext[j]);
  }
  SegDesc d{};  // one cell unless a grid below fits: every point then maps to cell 0
  d.cell = 1.0f;
  d.icell = 1.0f;
  d.g[0] = d.g[1] = d.g[2] = 1;
  if (!(n > 0.0) || !(emax > 0.0) || !(emax < 1e30)) {
    c0[c] = 1.0f;
    d0[c] = d;
    return;
  }
  for (int j = 0; j < 3; ++j) vol *= fmax(ext[j], emax * 1e-3);
  float e = (float)fmax(cbrt(vol / n * (double)K), emax * 1e-6);
  for (int it = 0; it < 64; ++it) {
    const float mg = e * 0.01f;
    double g0 = axis_cells(m[3], m[0] - mg, e), g1 = axis_cells(m[4], m[1] - mg, e), g2 = axis_cells(m[5], m[2] - mg, e);
    if (g0 * g1 * g2 <= (double)OCC_BITS) {
      d.mn[0] = m[0] - mg;
      d.mn[1] = m[1] - mg;
      d.mn[2] = m[2] - mg;
      d.cell = e;
      d.g[0] = (int)g0;
      d.g[1] = (int)g1;
      d.g[2] = (int)g2;
      break;
    }
    e *= 2.0f;
  }
  c0[c] = e;
  d.icell = __fdiv_rn(1.0f, d.cell);
  d0[c] = d;
}

// k-NN edge, step 2: occupied cells of the c0 grid (one bit per cell)
__global__ void seg_occ_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                          const uint32_t *__restrict__ off, int ncl, const SegDesc *__restrict__ d0,
                          uint8_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < off[ncl]) {
    const int c = cloud_of(off, ncl, i);
    const SegDesc d = d0[c];
    const uint32_t k = (uint32_t)cell_clamped(x[i], d.mn[0], d.cell, d.g[0]) +
                       (uint32_t)d.g[0] * ((uint32_t)cell_clamped(y[i], d.mn[1], d.cell, d.g[1]) +
                                           (uint32_t)d.g[1] * (uint32_t)cell_clamped(z[i], d.mn[2], d.cell, d.g[2]));
    uint8_t *f = flags + (size_t)c * OCC_BITS + k;  // racing plain stores of one value: no atomics
    if (!*f) *f = 1;
  }
}

// k-NN edge, step 3: occupied cells = set bits; surface density m / c0^2 from m points per occupied cell gives
// the k-th neighbour distance R = c0 sqrt(K / (pi m)); the edge is R / div.  Every thread of the block takes part
// and gets the result; sm holds a word per wave.
__device__ __forceinline__ float knn_cell_of(const uint32_t *__restrict__ off, int K, const float *__restrict__ c0,
                                             const uint32_t *__restrict__ words, double div, int c,
                                             uint32_t *sm) {
  uint32_t cnt = 0;
#pragma unroll 16
  for (int j = threadIdx.x; j < OCC_BITS / 4; j += blockDim.x) {
    uint32_t w = words[(size_t)c * (OCC_BITS / 4) + j];  // four cells: count the nonzero bytes
    w |= w >> 1;
    w |= w >> 2;
    w |= w >> 4;
    cnt += __popc(w & 0x01010101u);
  }
  cnt = wave_sum(cnt);
  if (lane() == 0) sm[threadIdx.x >> 6] = cnt;
  __syncthreads();
  uint32_t occ = 0;
  for (int i = 0; i < (int)(blockDim.x / WAVE); ++i) occ += sm[i];
  __syncthreads();  // sm is reused by the next call
  const double n = (double)(off[c + 1] - off[c]);
  const double m = n / (double)max(occ, 1u);
  const double e0 = (double)c0[c];
  const double R = e0 * sqrt((double)K / (3.14159265358979 * fmax(m, 1e-30)));
  return (float)fmin(fmax(R / div, e0 / 64.0), e0 * 2.0);
}

// step 3 with a block per cloud
__global__ __launch_bounds__(256) void seg_knn_cell_k(const uint32_t *__restrict__ off, int K,
                                                       const float *__restrict__ c0, const uint32_t *__restrict__ words,
                                                       double div, float *__restrict__ cell) {
  __shared__ uint32_t sm[4];
  const float e = knn_cell_of(off, K, c0, words, div, blockIdx.x, sm);
  if (threadIdx.x == 0) cell[blockIdx.x] = e;
}

// per cloud: origin (reference margin), dims, coarsened until the cloud's dense table fits its budget
__device__ __forceinline__ SegDesc seg_geom(const float *m, uint32_t n, float req, double per_point) {
  const double budget = fmax((double)SEG_MIN_CELLS, per_point * (double)n);
  SegDesc d{};
  const float mg = req * 0.01f;
  d.mn[0] = m[0] - mg;
  d.mn[1] = m[1] - mg;
  d.mn[2] = m[2] - mg;
  d.cell = req;
  d.g[0] = d.g[1] = d.g[2] = 1;
  const bool ok = n > 0 && m[0] <= m[3] && m[1] <= m[4] && m[2] <= m[5] && req > 0.0f;
  if (ok) {
    float e = req;
    for (int it = 0; it < 64 && e < FLT_MAX; ++it) {
      const double g0 = axis_cells(m[3], d.mn[0], e), g1 = axis_cells(m[4], d.mn[1], e), g2 = axis_cells(m[5], d.mn[2], e);
      if (g0 >= 1 && g1 >= 1 && g2 >= 1 && g0 * g1 * g2 <= budget) {
        d.cell = e;
        d.g[0] = (int)g0;
        d.g[1] = (int)g1;
        d.g[2] = (int)g2;
        break;
      }
      e *= 2.0f;
    }
  }
  d.icell = __fdiv_rn(1.0f, d.cell);
  return d;
}

__global__ void seg_geom_k(const float *__restrict__ mm, const uint32_t *__restrict__ off, int ncl,
                           const float *__restrict__ cell, double per_point, SegDesc *__restrict__ desc,
                           uint32_t *__restrict__ ncells) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncl) return;
  const SegDesc d = seg_geom(mm + 6 * c, off[c + 1] - off[c], cell[c], per_point);
  desc[c] = d;
  ncells[c] = (uint32_t)d.g[0] * (uint32_t)d.g[1] * (uint32_t)d.g[2];
}

// Up to LAYOUT_MAX clouds in one block: (from the integer bbox when mmu is given) each cloud's geometry, the
// exclusive scan of the clouds' cell counts into table offsets, and the key width; the work of
// seg_minmax_final_k, seg_geom_k, the scan and seg_toff_k in one launch.
constexpr int LAYOUT_MAX = 1024;
// With knn_words (at most KNN_FUSE_MAX clouds), the k-NN edges of step 3 are computed first by the whole block
// (into cell), instead of by seg_knn_cell_k.
constexpr int KNN_FUSE_MAX = 4;
__global__ __launch_bounds__(LAYOUT_MAX) void seg_layout_k(const uint32_t *__restrict__ mmu, float *__restrict__ mm,
                                                           const uint32_t *__restrict__ off, int ncl,
                                                           float *__restrict__ cell, double per_point,
                                                           SegDesc *__restrict__ desc, uint32_t *__restrict__ ncells,
                                                           uint32_t *__restrict__ toff,
                                                           const uint32_t *__restrict__ knn_words, int K,
                                                           const float *__restrict__ c0, double div) {
  __shared__ float cell_s[KNN_FUSE_MAX];
  if (knn_words) {
    __shared__ uint32_t sm[LAYOUT_MAX / WAVE];
    for (int cc = 0; cc < ncl; ++cc) {
      const float e = knn_cell_of(off, K, c0, knn_words, div, cc, sm);
      if (threadIdx.x == 0) {
        cell_s[cc] = e;
        cell[cc] = e;
      }
    }
    __syncthreads();
  }
  const int c = threadIdx.x;
  SegDesc d{};
  uint32_t nc = 0;
  if (c < ncl) {
    float m[6];
    for (int j = 0; j < 6; ++j) {
      m[j] = mmu ? ord2f(mmu[6 * c + j]) : mm[6 * c + j];
      if (mmu) mm[6 * c + j] = m[j];
    }
    d = seg_geom(m, off[c + 1] - off[c], knn_words ? cell_s[c] : cell[c], per_point);
    nc = (uint32_t)d.g[0] * (uint32_t)d.g[1] * (uint32_t)d.g[2];
    ncells[c] = nc;
  }
  // block-wide exclusive scan: inclusive per wave, then the wave totals
  __shared__ uint32_t wt[LAYOUT_MAX / WAVE];
  const uint32_t inc = wave_incl_scan(nc);
  const int w = c >> 6;
  if (lane() == WAVE - 1) wt[w] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int i = 0; i < LAYOUT_MAX / WAVE; ++i) {
    if (i < w) before += wt[i];
    total += wt[i];
  }
  if (c < ncl) {
    d.toff = before + inc - nc;
    desc[c] = d;
    toff[c] = d.toff;
  }
  if (c == 0) toff[ncl] = total;
}

__global__ void seg_toff_k(const uint32_t *__restrict__ toff, int ncl, SegDesc *__restrict__ desc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncl) desc[c].toff = toff[c];
}

// Counting sort by cell (cell order within a cloud, clouds in order): each point's cell key (its cloud's table
// offset plus the dense cell id) and its slot among the cell's points, taken with an atomic on the cell's count.
// The order of points within a cell is that of the atomics; no result depends on it (every neighbour list is
// ordered by (d2, index), every sum over a list follows that order), only which wave serves which query.
// Spatially ordered input puts runs of one cell on consecutive lanes: each run takes its slots with one atomic
// from its first lane (same-address atomics of a wave would serialise in the L2).
__global__ void seg_count_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                            const uint32_t *__restrict__ off, int ncl, const SegDesc *__restrict__ desc,
                            uint32_t *__restrict__ keys, uint32_t *__restrict__ slot, uint32_t *__restrict__ cnt) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  // the valid lanes of a wave are a prefix of it (i grows with the lane)
  if (i >= off[ncl]) return;
  const int c = cloud_of(off, ncl, i);
  const SegDesc d = desc[c];
  const uint32_t k = d.toff + (uint32_t)cell_clamped(x[i], d.mn[0], d.cell, d.g[0]) +
                     (uint32_t)d.g[0] * ((uint32_t)cell_clamped(y[i], d.mn[1], d.cell, d.g[1]) +
                                         (uint32_t)d.g[1] * (uint32_t)cell_clamped(z[i], d.mn[2], d.cell, d.g[2]));
  keys[i] = k;
  const int l = lane();
  // every wave-wide value is taken here, outside any branch: a ballot under a condition sees only the lanes
  // that take it
  const int active = popc(ballot(true));
  const uint32_t kprev = bperm(k, up_addr(1));
  const uint64_t heads = ballot(l == 0 || kprev != k);
  const uint64_t upto = (2ull << l) - 1ull;  // lanes 0..l (all 64 for l = 63)
  const int h = 63 - __clzll((long long)(heads & upto));  // first lane of this lane's run
  const uint64_t later = heads & ~upto;
  const int end = later ? __ffsll((unsigned long long)later) - 1 : active;
  uint32_t base = 0;
  if (l == h) base = atomicAdd(&cnt[k], (uint32_t)(end - h));
  slot[i] = bperm(base, h << 2) + (uint32_t)(l - h);
}

// after the scan cs[k] is the first cell-order position of cell k: point i goes to cs[key] + slot (cell-order
// SoA, perm = cell order -> batch position, rank = its inverse)
__global__ void seg_scatter_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                              const uint32_t *__restrict__ off, int ncl, const uint32_t *__restrict__ keys,
                              const uint32_t *__restrict__ slot, const uint32_t *__restrict__ cs,
                              float *__restrict__ xs, float *__restrict__ ys, float *__restrict__ zs,
                              uint32_t *__restrict__ perm, uint32_t *__restrict__ rank) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= off[ncl]) return;
  const uint32_t p = cs[keys[i]] + slot[i];
  xs[p] = x[i];
  ys[p] = y[i];
  zs[p] = z[i];
  perm[p] = i;
  rank[i] = p;
}

// the cell-count table cleared up to the batch's table size (toff[ncl]) plus the end entry
__global__ void seg_zero_table_k(uint32_t *__restrict__ cs, size_t bound, const uint32_t *__restrict__ total_cells) {
  const size_t m = min((size_t)*total_cells + 1, bound);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    cs[i] = 0u;
}

int SegGrid::build(const float *x, const float *y, const float *z, const uint32_t *off_, int ncl_, size_t n_bound_,
                 bool mode_knn, int K, const float *cell_dev, hipStream_t s) {
  ncl = ncl_;
  n_bound = n_bound_;
  RVP_CHECK(ncl > 0, RVP_E_ARG, "no clouds");
  RVP_CHECK(n_bound > 0, RVP_E_ARG, "n_bound must be positive");
  RVP_CHECK(n_bound < 0xfffffff0ull / 5, RVP_E_UNSUPPORTED, "batch too large");
  RVP_TRY(offs.alloc((ncl + 1) * sizeof(uint32_t), s));
  off = offs.as<uint32_t>();
  const unsigned cb = ceil_div(ncl, 64);
  RVP_TRY(mm.alloc(ncl * 6 * sizeof(float), s));
  RVP_TRY(desc.alloc(ncl * sizeof(SegDesc), s));
  RVP_TRY(cellv.alloc(ncl * sizeof(float), s));
  RVP_TRY(ncells.alloc((ncl + 1) * sizeof(uint32_t), s));
  RVP_TRY(toff.alloc((ncl + 1) * sizeof(uint32_t), s));
  Tmp mmu, bits;
  RVP_TRY(mmu.alloc(ncl * 6 * sizeof(uint32_t), s));
  const size_t bit_words = mode_knn ? (size_t)ncl * OCC_BITS / 4 : 0;
  if (mode_knn) RVP_TRY(bits.alloc((size_t)ncl * OCC_BITS, s));
  // offsets clamp, bbox identity and occupancy bytes cleared: one launch
  const unsigned pb = (unsigned)std::min<size_t>(
      1024, std::max<size_t>(1, ceil_div(std::max<size_t>(6 * (size_t)ncl, bit_words), 256)));
  RVP_LAUNCH("seg_prep_k", s, seg_prep_k<<<pb, 256, 0, s>>>(off_, ncl, (uint32_t)n_bound, offs.as<uint32_t>(),
                                                            mmu.as<uint32_t>(), bits.as<uint32_t>(), bit_words));
  const unsigned per = (unsigned)std::min<size_t>(64, std::max<size_t>(1, n_bound / ((size_t)ncl * 4096)));
  RVP_LAUNCH("seg_minmax_k", s, seg_minmax_k<<<dim3(per, ncl), 256, 0, s>>>(x, y, z, off, mmu.as<uint32_t>()));
  const float *cells = cell_dev;
  // k-NN edges of a few clouds computed by the layout launch (its inputs live until it is enqueued)
  Tmp c0, d0;
  const uint32_t *fuse_words = nullptr;
  const float *fuse_c0 = nullptr;
  if (mode_knn) {
    RVP_TRY(c0.alloc(ncl * sizeof(float), s));
    RVP_TRY(d0.alloc(ncl * sizeof(SegDesc), s));
    RVP_LAUNCH("seg_knn_c0_k", s, seg_knn_c0_k<<<cb, 64, 0, s>>>(mmu.as<uint32_t>(), mm.as<float>(), off, ncl, K,
                                                                 c0.as<float>(), d0.as<SegDesc>()));
    RVP_LAUNCH("seg_occ_k", s, seg_occ_k<<<ceil_div(n_bound, 256), 256, 0, s>>>(
                                   x, y, z, off, ncl, d0.as<SegDesc>(), bits.as<uint8_t>()));
    cells = cellv.as<float>();
    if (ncl > KNN_FUSE_MAX)
      RVP_LAUNCH("seg_knn_cell_k", s, seg_knn_cell_k<<<ncl, 256, 0, s>>>(off, K, c0.as<float>(), bits.as<uint32_t>(),
                                                                         tune_knn_cell_div(), cellv.as<float>()));
    else
      fuse_words = bits.as<uint32_t>(), fuse_c0 = c0.as<float>();
  }
  const double per_point = tune_seg_cells_per_point();
  if (ncl <= LAYOUT_MAX) {
    // the k-NN path has the float bbox already (seg_knn_c0_k); a few clouds' k-NN edges are computed here too
    RVP_LAUNCH("seg_layout_k", s, seg_layout_k<<<1, LAYOUT_MAX, 0, s>>>(
                                      mode_knn ? nullptr : mmu.as<uint32_t>(), mm.as<float>(), off, ncl,
                                      const_cast<float *>(cells), per_point, desc.as<SegDesc>(),
                                      ncells.as<uint32_t>(), toff.as<uint32_t>(), fuse_words, K, fuse_c0,
                                      tune_knn_cell_div()));
  } else {
    if (!mode_knn)
      RVP_LAUNCH("seg_minmax_final_k", s, seg_minmax_final_k<<<ceil_div(6 * ncl, 256), 256, 0, s>>>(
                                              ncl, mmu.as<uint32_t>(), mm.as<float>()));
    RVP_LAUNCH("seg_geom_k", s, seg_geom_k<<<cb, 64, 0, s>>>(mm.as<float>(), off, ncl, cells, per_point,
                                                             desc.as<SegDesc>(), ncells.as<uint32_t>()));
    RVP_TRY(exclusive_scan_u32(ncells.as<uint32_t>(), toff.as<uint32_t>(), ncl, toff.as<uint32_t>() + ncl, s));
    RVP_LAUNCH("seg_toff_k", s, seg_toff_k<<<cb, 64, 0, s>>>(toff.as<uint32_t>(), ncl, desc.as<SegDesc>()));
  }
  // the table holds at most sum(max(SEG_MIN_CELLS, per_point n_c)) cells
  const uint64_t table = (uint64_t)std::ceil(per_point * (double)n_bound) + (uint64_t)SEG_MIN_CELLS * (uint64_t)ncl;
  RVP_CHECK(table < 0xfff00000ull, RVP_E_UNSUPPORTED, "batch cell table too large");
  // counting sort by cell: counts (and each point's slot) in the table, one scan of the table turns the counts
  // into the cell starts (cstart[k] = first cell-order position of cell k, for k in [0, toff[ncl]]), scatter
  const uint32_t *total_cells = toff.as<uint32_t>() + ncl;
  const unsigned nb = ceil_div(n_bound, 256);
  Tmp keys, slot;
  RVP_TRY(cstart.alloc((table + 1) * sizeof(uint32_t), s));
  RVP_TRY(keys.alloc(n_bound * sizeof(uint32_t), s));
  RVP_TRY(slot.alloc(n_bound * sizeof(uint32_t), s));
  RVP_LAUNCH("seg_zero_table_k", s, seg_zero_table_k<<<(unsigned)std::min<size_t>(4096, ceil_div(table + 1, 256)), 256, 0, s>>>(
                                        cstart.as<uint32_t>(), table + 1, total_cells));
  RVP_LAUNCH("seg_count_k", s, seg_count_k<<<nb, 256, 0, s>>>(x, y, z, off, ncl, desc.as<SegDesc>(),
                                                              keys.as<uint32_t>(), slot.as<uint32_t>(),
                                                              cstart.as<uint32_t>()));
  RVP_TRY(exclusive_scan_u32_bounded(cstart.as<uint32_t>(), cstart.as<uint32_t>(), table + 1, total_cells, 1, nullptr,
                                     s));
  RVP_TRY(xs.alloc(n_bound * sizeof(float), s));
  RVP_TRY(ys.alloc(n_bound * sizeof(float), s));
  RVP_TRY(zs.alloc(n_bound * sizeof(float), s));
  RVP_TRY(perm.alloc(n_bound * sizeof(uint32_t), s));
  RVP_TRY(rank.alloc(n_bound * sizeof(uint32_t), s));
  RVP_LAUNCH("seg_scatter_k", s, seg_scatter_k<<<nb, 256, 0, s>>>(x, y, z, off, ncl, keys.as<uint32_t>(),
                                                                  slot.as<uint32_t>(), cstart.as<uint32_t>(),
                                                                  xs.as<float>(), ys.as<float>(), zs.as<float>(),
                                                                  perm.as<uint32_t>(), rank.as<uint32_t>()));
  return RVP_OK;
}

// ------------------------------------------------------------------------------------------------
// batched stages
// ------------------------------------------------------------------------------------------------
// S1 per cloud: k+1 nearest within the query's cloud, mean of sqrt of entries 1..k (sequential)
// Waves take runs of consecutive queries (cell order; run length from the device-side count, `forced` > 0
// fixes it); within a cloud, the previous query's k-th distance bounds the next one's (knn_bound_from).
// Only that k-th d2 is carried between queries (the previous position is re-read).  A wave's sums are
// deferred over its runs (DeferredMeans).  The launch has a wave per run or two: waves that finish early are
// replaced by new ones, which balances uneven query costs better than long-lived waves striding over runs.
__device__ __forceinline__ void sor_run(const SegArrays &A, int k, int h0, uint32_t run, uint32_t pw,
                                        uint32_t total, WaveLds &L, SumLds &S, DeferredMeans &dm,
                                        float *__restrict__ means) {
  const uint32_t q1 = min(run * pw + pw, total);
  float pd2 = -1.0f;  // k-th d2 of query q - 1 when it is in the run (< 0: none)
  for (uint32_t q = run * pw; q < q1; ++q) {
    const int c = cloud_of(A.off, A.ncl, q);
    const GridView g = seg_view(A, c);
    const float x = uniform(g.xs[q]), y = uniform(g.ys[q]), z = uniform(g.zs[q]);
    TopKT<uint32_t> t;
    t.init(k + 1);
    // the previous query of the run is q - 1, re-read from the cache; it bounds q only within one cloud
    knn_search<false>(g, x, y, z, t, L,
                      (pd2 >= 0.0f && q > A.off[c])
                          ? knn_bound_from(x, y, z, uniform(g.xs[q - 1]), uniform(g.ys[q - 1]), uniform(g.zs[q - 1]),
                                           pd2)
                          : __builtin_inff(),
                      h0);
    const int l = lane();
    const int m = min(t.cnt, k + 1);
    // entries beyond m add +0, which leaves the sum unchanged
    dm.add(S, (l >= 1 && l <= k && l < m) ? sqrtf(kd2(t.key)) : 0.0f, g.perm[q], k, means);
    pd2 = t.cnt >= t.K ? kd2(t.kth) : -1.0f;
  }
}

// WS (the work_steal knob) 1: persistent waves take runs from per-XCD ranges and steal (XcdWork), `chunk` runs
// at a time; 2: runs strided within per-XCD ranges (XcdStride); 0: runs strided over the launch's waves.  One
// instantiation per distribution, so each carries only its own loop's registers.
template <int WS>
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(8)
void seg_sor_means_k(SegArrays A, int k, int forced, int h0, float *__restrict__ means, uint32_t *__restrict__ queue,
                     uint32_t chunk) {
  __shared__ WaveLds lds[QW];
  __shared__ SumLds sums[QW];
  const int w = wave_in_block();
  const uint32_t total = A.off[A.ncl];
  const uint32_t pw = (uint32_t)knn_run_length(total, forced);
  const uint32_t nruns = (total + pw - 1) / pw;
  DeferredMeans dm;
  if constexpr (WS == 1) {
    XcdWork work;
    work.init(queue, nruns, chunk);
    uint32_t r0 = 0, r1 = 0;
    while (work.next(r0, r1))
      for (uint32_t run = r0; run < r1; ++run) sor_run(A, k, h0, run, pw, total, lds[w], sums[w], dm, means);
    work.finish();
  } else if constexpr (WS == 2) {  // XCD-contiguous ranges without atomics
    XcdStride it;
    it.init(nruns);
    uint32_t run = 0;
    while (it.next(run)) sor_run(A, k, h0, run, pw, total, lds[w], sums[w], dm, means);
  } else {
    for (uint32_t run = wave_id(); run < nruns; run += wave_count())
      sor_run(A, k, h0, run, pw, total, lds[w], sums[w], dm, means);
  }
  dm.flush(sums[w], k, means);
}

// H3 per cloud: capped radius lists (cloud-local indices), radius per cloud
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(8) void seg_radius_topk_k(SegArrays A, const float *__restrict__ radius, int K,
                                                         int per_wave, int32_t *__restrict__ idx_out,
                                                         float *__restrict__ d2_out, int32_t *__restrict__ counts) {
  __shared__ WaveLds lds[QW];
  __shared__ StageLdsSmall stage[QW];
  const int w = wave_in_block();
  WaveLds &L = lds[w];
  StageLdsSmall &S = stage[w];
  const uint32_t nq = A.off[A.ncl];
  const uint64_t runs = ((uint64_t)nq + (uint64_t)per_wave - 1) / (uint64_t)per_wave;
  const int l = lane();
  // runs of per_wave consecutive cell-order positions, strided over the launch's waves
  for (uint64_t run = wave_id(); run < runs; run += wave_count()) {
    const uint64_t q0 = run * (uint64_t)per_wave;
    const int nb = (int)min<uint64_t>((uint64_t)per_wave, (uint64_t)nq - q0);
    LaneQuery mine;
    int mc = 0;
    if (l < nb) {
      const uint32_t q = (uint32_t)q0 + (uint32_t)l;
      mc = cloud_of(A.off, A.ncl, q);
      const GridView g = seg_view(A, mc);
      mine.load(g, g.xs[q], g.ys[q], g.zs[q], g.perm[q], radius[mc]);
    }
    StagedWindow win;
    int32_t my_count = 0;
    for (int j = 0; j < nb; ++j) {
      const LaneQuery qj = mine.of_lane(j);
      const int c = __builtin_amdgcn_readlane(mc, j);
      const GridView g = seg_view(A, c);
      const float r = radius[c];
      const float r2 = __fmul_rn(r, r);
      TopK t;
      t.init(K);
      TopKSink<false> sink{t, L};
      win.ensure(g, qj.lo, qj.hi, c, L, S);
      if (win.total >= 0) scan_staged(S, win.total, qj.x, qj.y, qj.z, r2, sink);
      else scan_box(g, qj.x, qj.y, qj.z, qj.lo, qj.hi, false, qj.lo, qj.hi, r2, true, L, sink);
      topk_flush<false>(t, L);
      const int m = min(t.cnt, K);
      const uint32_t base = A.off[c];
      if (l < m) {
        if (idx_out) idx_out[(size_t)qj.row * K + l] = (int32_t)(key_idx(t.key) - base);
        if (d2_out) d2_out[(size_t)qj.row * K + l] = key_d2(t.key);
      }
      if (l == j) my_count = m;
    }
    if (l < nb) counts[mine.row] = my_count;
  }
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_batch_sor_exact(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                        size_t n_bound, int k, float alpha, float *out_aos, uint32_t *out_offsets, uint8_t *keep,
                        float *means, rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && out_aos && out_offsets, RVP_E_ARG, "null pointer");
  RVP_CHECK(nclouds > 0, RVP_E_ARG, "nclouds must be positive");
  RVP_CHECK(k >= 1 && k + 1 <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 63]");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  SegGrid G;
  RVP_TRY(G.build(x, y, z, offsets, nclouds, n_bound, true, k + 1, nullptr, s));
  Tmp mt, sums, vsum;
  float *mp = means;
  if (!mp) {
    RVP_TRY(mt.alloc(n_bound * sizeof(float), s));
    mp = mt.as<float>();
  }
  {
    const int forced = tune_queries_per_wave();
    unsigned blocks = query_blocks(n_bound);
    uint32_t *queue = nullptr, chunk = 1;
    const int mode = tune_work_steal();
    if (mode == 1) {
      // persistent: as many blocks as the device holds at once (fewer for small batches), runs in chunks of
      // about half a wave's share
      RVP_TRY(stream_work_queue(s, &queue));
      unsigned res = 1;
      RVP_TRY(resident_blocks(reinterpret_cast<const void *>(seg_sor_means_k<1>), QT, &res));
      const size_t runs = ceil_div(n_bound, (size_t)knn_run_length(n_bound, forced));
      blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(res, ceil_div(runs, QW)));
      chunk = (uint32_t)std::max<size_t>(1, runs / ((size_t)blocks * QW * 2));
      RVP_LAUNCH("seg_sor_means_k", s, seg_sor_means_k<1><<<blocks, QT, 0, s>>>(G.arrays(), k, forced, tune_knn_start_cells(), mp, queue,
                                                                               chunk));
    } else if (mode == 2) {
      RVP_LAUNCH("seg_sor_means_k", s, seg_sor_means_k<2><<<blocks, QT, 0, s>>>(G.arrays(), k, forced, tune_knn_start_cells(), mp, queue,
                                                                               chunk));
    } else {
      RVP_LAUNCH("seg_sor_means_k", s, seg_sor_means_k<0><<<blocks, QT, 0, s>>>(G.arrays(), k, forced, tune_knn_start_cells(), mp, queue,
                                                                               chunk));
    }
  }
  RVP_TRY(sums.alloc(nclouds * sizeof(float), s));
  RVP_TRY(vsum.alloc(nclouds * sizeof(float), s));
  // statistical_outlier_removal.cpp:44-51: the sum of the means, then of the squared deviations from their
  // mean, both sequential
  RVP_TRY(exact_seq_mean_var_seg(mp, G.off, nclouds, n_bound, sums.as<float>(), vsum.as<float>(), s));
  // keep flags, stable compaction over the whole batch (clouds stay in order) and each cloud's first output row
  return seg_keep_compact(x, y, z, mp, G.off, nclouds, n_bound, sums.as<float>(), vsum.as<float>(), alpha, keep,
                          out_aos, out_offsets, s);
}

int rvp_batch_radius_search(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                            size_t n_bound, const float *radius, const float *cell, int max_nn, int32_t *idx_out,
                            float *d2_out, int32_t *counts, rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && radius && counts, RVP_E_ARG, "null pointer");
  RVP_CHECK(nclouds > 0, RVP_E_ARG, "nclouds must be positive");
  RVP_CHECK(max_nn >= 1 && max_nn <= MAX_TOPK, RVP_E_UNSUPPORTED, "max_nn must be in [1, 64]");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  SegGrid G;
  RVP_TRY(G.build(x, y, z, offsets, nclouds, n_bound, false, 0, cell ? cell : radius, s));
  const int pw = queries_per_wave(n_bound);
  RVP_LAUNCH("seg_radius_topk_k", s, seg_radius_topk_k<<<query_blocks(ceil_div(n_bound, pw)), QT, 0, s>>>(
                                         G.arrays(), radius, max_nn, pw, idx_out, d2_out, counts));
  return RVP_OK;
}

}  // extern "C"
