// query.hip -- H3 radius search (capped top-K and CSR), H4 brute force, kNN, and the two SOR paths.
#include <algorithm>

#include "prims.hpp"
#include "query.hpp"

namespace rvp {

// ------------------------------------------------------------------------------------------------
// H3, capped: one wave per query keeps the max_nn smallest (d2, idx) within radius
// (spatial_hashing.cpp:165-223 with max_nn > 0)
// ------------------------------------------------------------------------------------------------
template <bool SELF>
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(8) void radius_topk_k(GridView g, const float *__restrict__ qx,
                                                     const float *__restrict__ qy, const float *__restrict__ qz,
                                                     uint32_t nq, int per_wave, float r, int K,
                                                     int32_t *__restrict__ idx_out, float *__restrict__ d2_out,
                                                     int32_t *__restrict__ counts, int32_t *__restrict__ total) {
  __shared__ WaveLds lds[QW];
  __shared__ StageLdsSmall stage[QW];
  const int w = wave_in_block();
  WaveLds &L = lds[w];
  StageLdsSmall &S = stage[w];
  // this wave's queries [q0, q0 + nb): lane j loads query q0 + j and computes its window once
  const uint64_t q0 = (uint64_t)(blockIdx.x * QW + w) * (uint64_t)per_wave;
  if (q0 >= nq) return;
  const int nb = (int)min<uint64_t>((uint64_t)per_wave, nq - q0);
  LaneQuery mine;
  if (lane() < nb) {
    const uint32_t q = (uint32_t)q0 + (uint32_t)lane();
    if (SELF) mine.load(g, g.xs[q], g.ys[q], g.zs[q], g.perm[q], r);
    else mine.load(g, qx[q], qy[q], qz[q], q, r);
  }
  const float r2 = __fmul_rn(r, r);
  int32_t my_count = 0, my_total = 0;
  StagedWindow win;
  for (int j = 0; j < nb; ++j) {
    const LaneQuery qj = mine.of_lane(j);
    TopK t;
    t.init(K);
    TopKSink<false> sink{t, L};
    win.ensure(g, qj.lo, qj.hi, 0, L, S);
    if (win.total >= 0) scan_staged(S, win.total, qj.x, qj.y, qj.z, r2, sink);
    else scan_box(g, qj.x, qj.y, qj.z, qj.lo, qj.hi, false, qj.lo, qj.hi, r2, true, L, sink);
    topk_flush<false>(t, L);
    const int m = min(t.cnt, K);
    const int l = lane();
    // whole rows: a row of max_nn words is written by one store per array, so the caches take full lines
    // without first reading them (entries past the count hold index -1 and d2 +inf)
    if (l < K) {
      if (idx_out) idx_out[(size_t)qj.row * K + l] = l < m ? (int32_t)key_idx(t.key) : -1;
      if (d2_out) d2_out[(size_t)qj.row * K + l] = l < m ? key_d2(t.key) : __builtin_inff();
    }
    if (l == j) {
      my_count = m;
      my_total = (int32_t)t.hits;
    }
  }
  if (lane() < nb) {
    counts[mine.row] = my_count;
    if (total) total[mine.row] = my_total;
  }
}

// ------------------------------------------------------------------------------------------------
// H3, unlimited: count, then fill CSR rows in enumeration order, then sort each row by (d2, idx)
// ------------------------------------------------------------------------------------------------
struct CountSink {
  uint32_t n = 0;
  __device__ __forceinline__ void operator()(bool, bool hit, uint64_t, uint32_t) { n += (uint32_t)popc(ballot(hit)); }
};

struct FillSink {
  uint64_t *out;  // keys of this row
  uint32_t *pos;  // payload (cell-order position), may be null
  uint32_t run = 0;
  __device__ __forceinline__ void operator()(bool, bool hit, uint64_t key, uint32_t p) {
    const uint64_t m = ballot(hit);
    if (hit) {
      const uint32_t at = run + (uint32_t)popc(m & lanemask_lt());
      out[at] = key;
      if (pos) pos[at] = p;
    }
    run += (uint32_t)popc(m);
  }
};

template <bool SELF>
__global__ __launch_bounds__(QT) void radius_count_k(GridView g, const float *__restrict__ qx,
                                                      const float *__restrict__ qy, const float *__restrict__ qz,
                                                      uint32_t nq, float r, int32_t *__restrict__ counts) {
  __shared__ WaveLds lds[QW];
  const int w = wave_in_block();
  const uint32_t q = blockIdx.x * QW + w;
  if (q >= nq) return;
  float x, y, z;
  uint32_t row;
  if (SELF) {
    x = g.xs[q]; y = g.ys[q]; z = g.zs[q];
    row = g.perm[q];
  } else {
    x = qx[q]; y = qy[q]; z = qz[q];
    row = q;
  }
  int lo[3], hi[3];
  window3(g, x, y, z, r, lo, hi);
  CountSink sink;
  scan_box(g, x, y, z, lo, hi, false, lo, hi, __fmul_rn(r, r), true, lds[w], sink);
  if (lane() == 0) counts[row] = (int32_t)sink.n;
}

template <bool SELF>
__global__ __launch_bounds__(QT) void radius_fill_k(GridView g, const float *__restrict__ qx,
                                                     const float *__restrict__ qy, const float *__restrict__ qz,
                                                     uint32_t nq, float r, const int64_t *__restrict__ offsets,
                                                     uint64_t *__restrict__ keys, uint32_t *__restrict__ pos) {
  __shared__ WaveLds lds[QW];
  const int w = wave_in_block();
  const uint32_t q = blockIdx.x * QW + w;
  if (q >= nq) return;
  float x, y, z;
  uint32_t row;
  if (SELF) {
    x = g.xs[q]; y = g.ys[q]; z = g.zs[q];
    row = g.perm[q];
  } else {
    x = qx[q]; y = qy[q]; z = qz[q];
    row = q;
  }
  int lo[3], hi[3];
  window3(g, x, y, z, r, lo, hi);
  const int64_t o = offsets[row];
  FillSink sink{keys + o, pos ? pos + o : nullptr};
  scan_box(g, x, y, z, lo, hi, false, lo, hi, __fmul_rn(r, r), true, lds[w], sink);
}

// rows of <= 64 keys: one wave each, registers only
__global__ __launch_bounds__(QT) void segsort_small_k(const int64_t *__restrict__ offsets, uint32_t nrows,
                                                       uint64_t *__restrict__ keys, uint32_t *__restrict__ pos) {
  const uint32_t r = blockIdx.x * QW + (uint32_t)wave_in_block();
  if (r >= nrows) return;
  const int64_t b = offsets[r];
  const int m = (int)(offsets[r + 1] - b);
  if (m <= 1 || m > WAVE) return;
  const int l = lane();
  uint64_t k = l < m ? keys[b + l] : ~0ull;
  uint32_t p = (pos && l < m) ? pos[b + l] : 0u;
  if (pos) bitonic_sort64<true>(k, p);
  else bitonic_sort64<false>(k, p);
  if (l < m) {
    keys[b + l] = k;
    if (pos) pos[b + l] = p;
  }
}

// rows longer than 64: one block per row, mirror-bitonic network with virtual +inf padding (every
// compare-exchange keeps the minimum at the lower index, so padded slots never move into the row);
// rows up to LDS_SORT elements are sorted in LDS, longer ones in place in global memory.
constexpr int LDS_SORT = 4096;

__device__ __forceinline__ void cx_mem(uint64_t *k, uint32_t *p, uint32_t a, uint32_t b, uint32_t m) {
  if (b >= m) return;
  const uint64_t ka = k[a], kb = k[b];
  if (kb < ka) {
    k[a] = kb;
    k[b] = ka;
    if (p) {
      const uint32_t t = p[a];
      p[a] = p[b];
      p[b] = t;
    }
  }
}

__device__ void bitonic_block(uint64_t *k, uint32_t *p, uint32_t m, bool shared_mem) {
  uint32_t P = 1;
  while (P < m) P <<= 1;
  for (uint32_t s = 2; s <= P; s <<= 1) {
    for (uint32_t j = s >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P / 2; i += blockDim.x) {
        const uint32_t a = (i / j) * 2 * j + (i % j);
        const uint32_t b = (j == (s >> 1)) ? (a / s) * s + (s - 1 - (a % s)) : a + j;
        cx_mem(k, p, a, b, m);
      }
      if (shared_mem) __syncthreads();
      else {
        __threadfence_block();
        __syncthreads();
      }
    }
  }
}

__global__ __launch_bounds__(256) void segsort_large_k(const int64_t *__restrict__ offsets, uint32_t nrows,
                                                        uint64_t *__restrict__ keys, uint32_t *__restrict__ pos) {
  __shared__ uint64_t sk[LDS_SORT];
  __shared__ uint32_t sp[LDS_SORT];
  for (uint32_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    const int64_t b = offsets[r];
    const uint32_t m = (uint32_t)(offsets[r + 1] - b);
    if (m <= WAVE) continue;
    if (m <= LDS_SORT) {
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        sk[i] = keys[b + i];
        if (pos) sp[i] = pos[b + i];
      }
      __syncthreads();
      bitonic_block(sk, pos ? sp : nullptr, m, true);
      for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
        keys[b + i] = sk[i];
        if (pos) pos[b + i] = sp[i];
      }
      __syncthreads();
    } else {
      bitonic_block(keys + b, pos ? pos + b : nullptr, m, false);
    }
  }
}

__global__ void unpack_rows_k(const uint64_t *__restrict__ keys, size_t total, int32_t *__restrict__ idx,
                              float *__restrict__ d2) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const uint64_t k = keys[i];
  if (idx) idx[i] = (int32_t)key_idx(k);
  if (d2) d2[i] = key_d2(k);
}

// sorted CSR rows -> fixed rows of stride K (first min(len, K) entries)
__global__ void truncate_rows_k(const uint64_t *__restrict__ keys, const int64_t *__restrict__ offsets,
                                uint32_t nrows, int K, int32_t *__restrict__ idx, float *__restrict__ d2,
                                int32_t *__restrict__ counts, int32_t *__restrict__ total) {
  const uint32_t r = blockIdx.x;
  if (r >= nrows) return;
  const int64_t b = offsets[r];
  const int64_t m = offsets[r + 1] - b;
  const int keep = (int)std::min<int64_t>(m, K);
  for (int j = threadIdx.x; j < keep; j += blockDim.x) {
    const uint64_t k = keys[b + j];
    if (idx) idx[(size_t)r * K + j] = (int32_t)key_idx(k);
    if (d2) d2[(size_t)r * K + j] = key_d2(k);
  }
  if (threadIdx.x == 0) {
    counts[r] = keep;
    if (total) total[r] = (int32_t)m;
  }
}

// ------------------------------------------------------------------------------------------------
// H4: radius_search_sc / _rvv (radius_search.cpp:12-59): first max_nn hits in index order
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(QT) void bruteforce_k(const float *__restrict__ x, const float *__restrict__ y,
                                                    const float *__restrict__ z, uint32_t n,
                                                    const float *__restrict__ qx, const float *__restrict__ qy,
                                                    const float *__restrict__ qz, uint32_t nq, float r, int max_nn,
                                                    int32_t *__restrict__ idx_out, float *__restrict__ d2_out,
                                                    int32_t *__restrict__ counts) {
  const uint32_t q = blockIdx.x * QW + (uint32_t)wave_in_block();
  if (q >= nq) return;
  const int l = lane();
  const float a = qx[q], b = qy[q], c = qz[q];
  const float r2 = __fmul_rn(r, r);
  uint32_t cnt = 0;
  for (uint32_t base = 0; base < n; base += WAVE) {
    const uint32_t j = base + (uint32_t)l;
    bool hit = false;
    float d2 = 0.f;
    if (j < n) {
      d2 = dist2(x[j], y[j], z[j], a, b, c);
      hit = d2 <= r2;
    }
    const uint64_t m = ballot(hit);
    if (hit) {
      const uint32_t at = cnt + (uint32_t)popc(m & lanemask_lt());
      if (at < (uint32_t)max_nn) {
        idx_out[(size_t)q * max_nn + at] = (int32_t)j;
        if (d2_out) d2_out[(size_t)q * max_nn + at] = d2;
      }
    }
    cnt += (uint32_t)popc(m);
  }
  if (l == 0) counts[q] = (int32_t)std::min<uint32_t>(cnt, (uint32_t)max_nn);
}

// ------------------------------------------------------------------------------------------------
// k nearest (rows of stride k, padded with -1 / +inf)
// ------------------------------------------------------------------------------------------------
template <bool SELF>
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(8)
void knn_k(GridView g, const float *__restrict__ qx, const float *__restrict__ qy, const float *__restrict__ qz,
           uint32_t nq, int k, int32_t *__restrict__ idx_out, float *__restrict__ d2_out) {
  __shared__ WaveLds lds[QW];
  const int w = wave_in_block();
  const uint32_t q = blockIdx.x * QW + w;
  if (q >= nq) return;
  float x, y, z;
  uint32_t row;
  if (SELF) {
    x = g.xs[q]; y = g.ys[q]; z = g.zs[q];
    row = g.perm[q];
  } else {
    x = qx[q]; y = qy[q]; z = qz[q];
    row = q;
  }
  TopK t;
  t.init(k);
  knn_search<false>(g, x, y, z, t, lds[w]);
  const int l = lane();
  if (l < k) {
    const bool ok = l < t.cnt;
    if (idx_out) idx_out[(size_t)row * k + l] = ok ? (int32_t)key_idx(t.key) : -1;
    if (d2_out) d2_out[(size_t)row * k + l] = ok ? key_d2(t.key) : __uint_as_float(0x7f800000u);
  }
}

// ------------------------------------------------------------------------------------------------
// SOR: per-point mean neighbour distance, then the reference's sequential fp32 global statistics
// ------------------------------------------------------------------------------------------------
// S1 (statistical_outlier_removal.cpp:27-41): k+1 smallest d2 incl. self; sum of sqrt of entries 1..k in
// ascending order, sequentially; divide by k.
// Waves take runs of per_wave consecutive queries (cell order), strided over the launch: from the second
// query of a run on, the previous query's k-th distance bounds the current one (knn_bound_from), so the first
// box is already the certified one and only hits within the bound are staged.  The sums of a wave's queries
// are deferred (DeferredMeans); the launch has one wave per run.
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(8)
void sor_knn_means_k(GridView g, int k, int per_wave, int h0, float *__restrict__ means) {
  __shared__ WaveLds lds[QW];
  __shared__ SumLds sums[QW];
  const int w = wave_in_block();
  const uint32_t nruns = (g.n + (uint32_t)per_wave - 1) / (uint32_t)per_wave;
  DeferredMeans dm;
  for (uint32_t run = blockIdx.x * QW + (uint32_t)w; run < nruns; run += gridDim.x * QW) {
    const uint32_t q0 = run * (uint32_t)per_wave, q1 = min(q0 + (uint32_t)per_wave, g.n);
    float px = 0.f, py = 0.f, pz = 0.f, pd2 = -1.0f;  // previous query and its k-th d2 (< 0: none)
    for (uint32_t q = q0; q < q1; ++q) {
      const float x = uniform(g.xs[q]), y = uniform(g.ys[q]), z = uniform(g.zs[q]);
      TopKT<uint32_t> t;  // distances only: the means read the multiset of the k+1 smallest
      t.init(k + 1);
      knn_search<false>(g, x, y, z, t, lds[w], knn_bound_from(x, y, z, px, py, pz, pd2), h0);
      const int l = lane();
      const int m = min(t.cnt, k + 1);
      // entries beyond m add +0, which leaves the sum unchanged
      dm.add(sums[w], (l >= 1 && l <= k && l < m) ? sqrtf(kd2(t.key)) : 0.0f, g.perm[q], k, means);
      px = x;
      py = y;
      pz = z;
      pd2 = t.cnt >= t.K ? kd2(t.kth) : -1.0f;
    }
  }
  dm.flush(sums[w], k, means);
}

// S2 (statistical_outlier_removal.cpp:246-269): radius list capped at k+1; mean over min(k, m-1) entries
// after the first; search_radius when m <= 1.
__global__ __launch_bounds__(QT) void sor_radius_means_k(GridView g, int k, float radius,
                                                          float *__restrict__ means) {
  __shared__ WaveLds lds[QW];
  const int w = wave_in_block();
  const uint32_t q = blockIdx.x * QW + w;
  if (q >= g.n) return;
  WaveLds &L = lds[w];
  const float x = g.xs[q], y = g.ys[q], z = g.zs[q];
  int lo[3], hi[3];
  window3(g, x, y, z, radius, lo, hi);
  TopKT<uint32_t> t;
  t.init(k + 1);
  TopKSink<false, uint32_t> sink{t, L};
  scan_box(g, x, y, z, lo, hi, false, lo, hi, __fmul_rn(radius, radius), true, L, sink);
  topk_flush<false>(t, L);
  const int l = lane();
  const int m = min(t.cnt, k + 1);
  float mean = radius;
  if (m > 1) {
    const int valid_k = min(k, m - 1);
    const float sq = (l >= 1 && l <= valid_k) ? sqrtf(kd2(t.key)) : 0.0f;
    float s = 0.0f;
    for (int j = 1; j <= valid_k; ++j) s = __fadd_rn(s, readlanef(sq, j));
    mean = __fdiv_rn(s, __int2float_rn(valid_k));
  }
  if (l == 0) means[g.perm[q]] = mean;
}

// Global statistics exactly as filter_by_mean_dists / sor_sc (statistical_outlier_removal.cpp:43-55,
// 193-206): sequential fp32 sums in index order (exact_seq_sum), then mean = sum / n,
// std = sqrt(varsum / n), threshold = mean + alpha * std.  stats = {mean, std, threshold, sum, varsum}.
__global__ void sor_mean_k(uint32_t n, float *__restrict__ stats) {
  if (threadIdx.x == 0) stats[0] = stats[3] / __uint2float_rn(n);
}

__global__ void sor_threshold_k(uint32_t n, float alpha, float *__restrict__ stats) {
  if (threadIdx.x != 0) return;
  const float gstd = sqrtf(stats[4] / __uint2float_rn(n));
  stats[1] = gstd;
  stats[2] = stats[0] + alpha * gstd;
}

__global__ void sor_keep_k(const float *__restrict__ means, uint32_t n, const float *__restrict__ stats,
                           uint8_t *__restrict__ keep) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keep[i] = means[i] <= stats[2] ? 1 : 0;
}

static int sor_tail(const Grid &grid, const float *x, const float *y, const float *z, size_t n, float alpha,
                    float *means, float *out_aos, size_t *n_out, uint8_t *keep, hipStream_t s) {
  Tmp stats, keep_t, cnt;
  RVP_TRY(stats.alloc(5 * sizeof(float), s));
  float *st = stats.as<float>();
  RVP_TRY(exact_seq_sum(means, n, 0, nullptr, st + 3, s));
  RVP_LAUNCH("sor_mean_k", s, sor_mean_k<<<1, 64, 0, s>>>((uint32_t)n, st));
  RVP_TRY(exact_seq_sum(means, n, 1, st, st + 4, s));
  RVP_LAUNCH("sor_threshold_k", s, sor_threshold_k<<<1, 64, 0, s>>>((uint32_t)n, alpha, st));
  uint8_t *kp = keep;
  if (!kp) {
    RVP_TRY(keep_t.alloc(n, s));
    kp = keep_t.as<uint8_t>();
  }
  RVP_LAUNCH("sor_keep_k", s, sor_keep_k<<<ceil_div(n, 256), 256, 0, s>>>(means, (uint32_t)n, stats.as<float>(), kp));
  RVP_TRY(cnt.alloc(sizeof(uint32_t), s));
  RVP_TRY(compact_xyz_aos(x, y, z, kp, 1, n, out_aos, cnt.as<uint32_t>(), s));
  uint32_t m = 0;
  RVP_HIP(hipMemcpyAsync(&m, cnt.p, sizeof(m), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *n_out = m;
  (void)grid;
  return RVP_OK;
}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------
static int check_built(const rvp_hash *h) {
  RVP_CHECK(h, RVP_E_ARG, "null hash");
  RVP_CHECK(h->g.built, RVP_E_STATE, "hash not built");
  return RVP_OK;
}

int radius_csr(const Grid &G, const float *qx, const float *qy, const float *qz, size_t nq, float r,
                      const int64_t *offsets, uint64_t *keys, uint32_t *pos, hipStream_t s) {
  const GridView g = G.view();
  const bool self = qx == nullptr;
  const unsigned nb = ceil_div(nq, QW);
  if (self) RVP_LAUNCH("radius_fill_k", s, radius_fill_k<true><<<nb, QT, 0, s>>>(g, qx, qy, qz, (uint32_t)nq, r, offsets, keys, pos));
  else RVP_LAUNCH("radius_fill_k", s, radius_fill_k<false><<<nb, QT, 0, s>>>(g, qx, qy, qz, (uint32_t)nq, r, offsets, keys, pos));
  RVP_LAUNCH("segsort_small_k", s, segsort_small_k<<<nb, QT, 0, s>>>(offsets, (uint32_t)nq, keys, pos));
  RVP_LAUNCH("segsort_large_k", s, segsort_large_k<<<std::min<unsigned>(ceil_div(nq, 1), 4096), 256, 0, s>>>(offsets, (uint32_t)nq, keys, pos));
  return RVP_OK;
}

static int radius_counts(const Grid &G, const float *qx, const float *qy, const float *qz, size_t nq, float r,
                         int32_t *counts, hipStream_t s) {
  const GridView g = G.view();
  const unsigned nb = ceil_div(nq, QW);
  if (!qx) RVP_LAUNCH("radius_count_k", s, radius_count_k<true><<<nb, QT, 0, s>>>(g, qx, qy, qz, (uint32_t)nq, r, counts));
  else RVP_LAUNCH("radius_count_k", s, radius_count_k<false><<<nb, QT, 0, s>>>(g, qx, qy, qz, (uint32_t)nq, r, counts));
  return RVP_OK;
}

static int query_args(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t &nq) {
  RVP_TRY(check_built(h));
  if (!qx) {
    RVP_CHECK(!qy && !qz, RVP_E_ARG, "self queries: pass qx = qy = qz = NULL");
    nq = h->g.n;
  } else {
    RVP_CHECK(qy && qz, RVP_E_ARG, "null query pointer");
  }
  RVP_CHECK(nq < 0x7fffffffull * QW, RVP_E_UNSUPPORTED, "too many queries");
  return RVP_OK;
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_radius_search(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq, float radius,
                      int max_nn, int32_t *idx_out, float *d2_out, int32_t *counts, int32_t *total, rvp_stream_t st) {
  RVP_TRY(query_args(h, qx, qy, qz, nq));
  RVP_CHECK(max_nn > 0, RVP_E_ARG, "rvp_radius_search needs max_nn > 0 (use rvp_radius_count/fill for unlimited)");
  RVP_CHECK(counts, RVP_E_ARG, "null counts");
  RVP_CHECK(radius >= 0.0f, RVP_E_ARG, "negative radius");
  if (nq == 0) return RVP_OK;
  hipStream_t s = S(st);
  const GridView g = h->g.view();
  if (max_nn <= MAX_TOPK) {
    const int pw = queries_per_wave(nq);
    const unsigned nb = ceil_div(ceil_div(nq, pw), QW);
    if (!qx) RVP_LAUNCH("radius_topk_k", s, radius_topk_k<true><<<nb, QT, 0, s>>>(g, qx, qy, qz, (uint32_t)nq, pw, radius, max_nn, idx_out, d2_out,
                                                   counts, total));
    else RVP_LAUNCH("radius_topk_k", s, radius_topk_k<false><<<nb, QT, 0, s>>>(g, qx, qy, qz, (uint32_t)nq, pw, radius, max_nn, idx_out, d2_out,
                                                counts, total));
    RVP_LAUNCHED();
  } else {
    Tmp cnt, off, keys;
    RVP_TRY(cnt.alloc(nq * sizeof(int32_t), s));
    RVP_TRY(off.alloc((nq + 1) * sizeof(int64_t), s));
    RVP_TRY(radius_counts(h->g, qx, qy, qz, nq, radius, cnt.as<int32_t>(), s));
    int64_t tot = 0;
    RVP_TRY(rvp_exclusive_scan(cnt.as<int32_t>(), nq, off.as<int64_t>(), &tot, st));
    RVP_TRY(keys.alloc((size_t)tot * sizeof(uint64_t), s));
    RVP_TRY(radius_csr(h->g, qx, qy, qz, nq, radius, off.as<int64_t>(), keys.as<uint64_t>(), nullptr, s));
    // CSR rows are indexed by output row already (self mode writes at the original index)
    RVP_LAUNCH("truncate_rows_k", s, truncate_rows_k<<<(unsigned)nq, 64, 0, s>>>(keys.as<uint64_t>(), off.as<int64_t>(), (uint32_t)nq, max_nn,
                                                idx_out, d2_out, counts, total));
    RVP_LAUNCHED();
  }
  RVP_HIP(hipStreamSynchronize(s));
  return RVP_OK;
}

int rvp_radius_count(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq, float radius,
                     int32_t *counts, rvp_stream_t st) {
  RVP_TRY(query_args(h, qx, qy, qz, nq));
  RVP_CHECK(counts, RVP_E_ARG, "null counts");
  if (nq == 0) return RVP_OK;
  RVP_TRY(radius_counts(h->g, qx, qy, qz, nq, radius, counts, S(st)));
  RVP_HIP(hipStreamSynchronize(S(st)));
  return RVP_OK;
}

int rvp_radius_fill(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq, float radius,
                    const int64_t *offsets, int32_t *idx_out, float *d2_out, rvp_stream_t st) {
  RVP_TRY(query_args(h, qx, qy, qz, nq));
  RVP_CHECK(offsets, RVP_E_ARG, "null offsets");
  if (nq == 0) return RVP_OK;
  hipStream_t s = S(st);
  int64_t tot = 0;
  RVP_HIP(hipMemcpyAsync(&tot, offsets + nq, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  Tmp keys;
  RVP_TRY(keys.alloc((size_t)tot * sizeof(uint64_t), s));
  RVP_TRY(radius_csr(h->g, qx, qy, qz, nq, radius, offsets, keys.as<uint64_t>(), nullptr, s));
  if (tot > 0) {
    RVP_LAUNCH("unpack_rows_k", s, unpack_rows_k<<<ceil_div((size_t)tot, 256), 256, 0, s>>>(keys.as<uint64_t>(), (size_t)tot, idx_out, d2_out));
  }
  RVP_HIP(hipStreamSynchronize(s));
  return RVP_OK;
}

int rvp_radius_bruteforce(const float *x, const float *y, const float *z, size_t n, const float *qx, const float *qy,
                          const float *qz, size_t nq, float radius, int max_nn, int32_t *idx_out, float *d2_out,
                          int32_t *counts, rvp_stream_t st) {
  RVP_CHECK(max_nn >= 0, RVP_E_ARG, "max_nn must be >= 0");
  RVP_CHECK(counts && (idx_out || max_nn == 0), RVP_E_ARG, "null output");
  RVP_CHECK(n < 0xffffffffull && nq < 0x7fffffffull, RVP_E_UNSUPPORTED, "size");
  if (nq == 0) return RVP_OK;
  RVP_TRY(pool_init());
  RVP_LAUNCH("bruteforce_k", S(st), bruteforce_k<<<ceil_div(nq, QW), QT, 0, S(st)>>>(x, y, z, (uint32_t)n, qx, qy, qz, (uint32_t)nq, radius, max_nn,
                                                   idx_out, d2_out, counts));
  RVP_LAUNCHED();
  RVP_HIP(hipStreamSynchronize(S(st)));
  return RVP_OK;
}

int rvp_knn(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq, int k, int32_t *idx_out,
            float *d2_out, rvp_stream_t st) {
  RVP_TRY(query_args(h, qx, qy, qz, nq));
  RVP_CHECK(k >= 1 && k <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 64]");
  if (nq == 0) return RVP_OK;
  const GridView g = h->g.view();
  const unsigned nb = ceil_div(nq, QW);
  if (!qx) RVP_LAUNCH("knn_k", S(st), knn_k<true><<<nb, QT, 0, S(st)>>>(g, qx, qy, qz, (uint32_t)nq, k, idx_out, d2_out));
  else RVP_LAUNCH("knn_k", S(st), knn_k<false><<<nb, QT, 0, S(st)>>>(g, qx, qy, qz, (uint32_t)nq, k, idx_out, d2_out));
  RVP_HIP(hipStreamSynchronize(S(st)));
  return RVP_OK;
}

int rvp_sor_exact(const float *x, const float *y, const float *z, size_t n, int k, float alpha, float *out_aos,
                  size_t *n_out, uint8_t *keep, float *means, rvp_stream_t st) {
  RVP_CHECK(n_out, RVP_E_ARG, "null n_out");
  RVP_CHECK(k >= 1 && k + 1 <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 63]");
  *n_out = 0;
  if (n == 0) return RVP_OK;
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  Grid G;
  RVP_TRY(G.build_for_knn(x, y, z, n, k + 1, s));
  Tmp mt;
  float *mp = means;
  if (!mp) {
    RVP_TRY(mt.alloc(n * sizeof(float), s));
    mp = mt.as<float>();
  }
  const int pw = knn_run_length(n, tune_queries_per_wave());
  RVP_LAUNCH("sor_knn_means_k", s, sor_knn_means_k<<<ceil_div(ceil_div(n, pw), QW), QT, 0, s>>>(G.view(), k, pw,
                                                                                          tune_knn_start_cells(), mp));
  return sor_tail(G, x, y, z, n, alpha, mp, out_aos, n_out, keep, s);
}

int rvp_sor_radius(const rvp_hash *h, int k, float alpha, float search_radius, float *out_aos, size_t *n_out,
                   uint8_t *keep, float *means, rvp_stream_t st) {
  RVP_TRY(check_built(h));
  RVP_CHECK(n_out, RVP_E_ARG, "null n_out");
  RVP_CHECK(k >= 1 && k + 1 <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 63]");
  *n_out = 0;
  const size_t n = h->g.n;
  if (n == 0) return RVP_OK;
  hipStream_t s = S(st);
  Tmp mt;
  float *mp = means;
  if (!mp) {
    RVP_TRY(mt.alloc(n * sizeof(float), s));
    mp = mt.as<float>();
  }
  RVP_LAUNCH("sor_radius_means_k", s, sor_radius_means_k<<<ceil_div(n, QW), QT, 0, s>>>(h->g.view(), k, search_radius, mp));
  return sor_tail(h->g, h->g.x, h->g.y, h->g.z, n, alpha, mp, out_aos, n_out, keep, s);
}

}  // extern "C"
