// query.hpp -- wave-level neighbour machinery shared by the radius, kNN, SOR and normals kernels.
//
// One wavefront (64 lanes) serves one query.  Candidates are the points of a box of cells; because the
// linear cell id is x-fastest, each (y, z) row of the box is ONE contiguous slice of the cell-ordered
// arrays, so a 3x3x3 box is 9 coalesced slices.  The 64 lanes walk the concatenated slices 64 at a time,
// compute D1 exactly as the reference (dist2), and offer hits to a per-wave top-K: hits are compacted
// with a ballot prefix into an LDS staging buffer and, when it fills, bitonic-sorted across the lanes
// and merged into the sorted best list held one entry per lane.  Keys are (d2 bits << 32 | index), so
// the order is the reference's (d2, idx) pair order (spatial_hashing.cpp:203-209).
#pragma once

#include <algorithm>

#include "grid.hpp"

namespace rvp {

// Occupancy floor for a query kernel: the compiler keeps each wave within the registers that let n waves share
// a SIMD, spilling scalar registers beyond that into lanes of vector registers.  The search kernels are bound by
// instruction issue, and the extra resident waves hide more stalls than the spill moves cost (k-NN SOR means at
// 8 waves instead of 7: 0.182 -> 0.160 ms on the LiDAR frame).  A kernel whose vector registers cannot fit is left
// as it is by the compiler, hence 7 for the radius-normals variants (72 VGPRs).
#define RVP_WAVES_PER_SIMD(n) __attribute__((amdgpu_waves_per_eu(n)))

constexpr int QT = 256;         // threads per query block
constexpr int QW = QT / WAVE;   // waves (queries in flight) per block
constexpr int MAX_TOPK = 64;    // entries one wave keeps in registers

struct WaveLds {
  uint64_t bk[WAVE];   // staged hit keys
  uint32_t excl[WAVE]; // row-slice prefix offsets
  uint32_t s0[WAVE], l0[WAVE], s1[WAVE];
};

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// KT = uint64_t: key (d2 bits << 32 | idx), the reference's (d2, idx) order; KT = uint32_t: d2 bits only, for
// callers that read nothing but the multiset of the K smallest distances (SOR's means)
template <class KT> struct TopKT {
  KT key;        // this lane's entry; lanes [0, cnt) sorted ascending, the rest all-ones
  uint32_t pay;
  int cnt;       // valid entries (<= 64)
  int K;
  KT kth;        // key at K-1 once cnt >= K, else `bound` (hits must beat it to be staged)
  KT bound;      // a key known to exceed the K-th smallest (all-ones: none known)
  int nb;        // staged entries
  uint32_t hits; // all hits offered (untruncated count)
  __device__ __forceinline__ void init(int k) {
    key = ~(KT)0;
    pay = 0;
    cnt = 0;
    K = k;
    kth = ~(KT)0;
    bound = ~(KT)0;
    nb = 0;
    hits = 0;
  }
};
using TopK = TopKT<uint64_t>;

__device__ __forceinline__ uint64_t read_lane(uint64_t v, int l) { return readlane64(v, l); }
__device__ __forceinline__ uint32_t read_lane(uint32_t v, int l) { return readlaneu(v, l); }
__device__ __forceinline__ float kd2(uint64_t k) { return key_d2(k); }
__device__ __forceinline__ float kd2(uint32_t k) { return __uint_as_float(k); }
// the TopK key of a packed (d2, idx) key
template <class KT> __device__ __forceinline__ KT as_key(uint64_t k) { return (KT)(k >> (64 - 8 * sizeof(KT))); }

// lanes whose lane index has bit b set (b a power of two below 64)
__host__ __device__ constexpr uint64_t lanes_with(int b) {
  return b == 1    ? 0xAAAAAAAAAAAAAAAAull
         : b == 2  ? 0xCCCCCCCCCCCCCCCCull
         : b == 4  ? 0xF0F0F0F0F0F0F0F0ull
         : b == 8  ? 0xFF00FF00FF00FF00ull
         : b == 16 ? 0xFFFF0000FFFF0000ull
                   : 0xFFFFFFFF00000000ull;
}

// Compare-exchange with lane l ^ m: lanes of `upper` keep the larger key, the others the smaller ("mirror"
// bitonic formulation).  The comparison is one wave-wide mask, flipped for the upper lanes by a scalar xor.
// Without a payload, equal keys may be taken from either side (they are the same value); with one, each
// side keeps its own entry on a tie.
template <bool PAY, class KT>
__device__ __forceinline__ void cx(KT &k, uint32_t &p, int m, uint64_t upper) {
  const int a = xor_addr(m);
  const KT ok = bperm(k, a);
  uint64_t take;
  if (PAY) take = (ballot(ok < k) & ~upper) | (ballot(ok > k) & upper);
  else take = ballot(ok < k) ^ upper;
  const uint32_t op = PAY ? bperm(p, a) : 0u;
  if (__builtin_amdgcn_inverse_ballot_w64(take)) {
    k = ok;
    if (PAY) p = op;
  }
}

constexpr int log2i(int v) { return v <= 1 ? 0 : 1 + log2i(v >> 1); }

// sorts every group of SIZE lanes ascending (counted loops, so that both unroll completely)
template <bool PAY, int SIZE, class KT> __device__ __forceinline__ void bitonic_sort_n(KT &k, uint32_t &p) {
#pragma unroll
  for (int a = 1; a <= log2i(SIZE); ++a) {
    const int s = 1 << a;
    cx<PAY>(k, p, s - 1, lanes_with(s >> 1));
#pragma unroll
    for (int b = a - 2; b >= 0; --b) cx<PAY>(k, p, 1 << b, lanes_with(1 << b));
  }
}

// sorts a bitonic sequence held in each group of SIZE lanes
template <bool PAY, int SIZE, class KT> __device__ __forceinline__ void bitonic_clean_n(KT &k, uint32_t &p) {
#pragma unroll
  for (int b = log2i(SIZE) - 1; b >= 0; --b) cx<PAY>(k, p, 1 << b, lanes_with(1 << b));
}

// sorts every group of `size` lanes ascending (size: power of two <= 64, wave-uniform)
template <bool PAY, class KT> __device__ __forceinline__ void bitonic_sort_groups(KT &k, uint32_t &p, int size) {
  if (size > 32) bitonic_sort_n<PAY, 64>(k, p);
  else if (size > 16) bitonic_sort_n<PAY, 32>(k, p);
  else if (size > 8) bitonic_sort_n<PAY, 16>(k, p);
  else if (size > 4) bitonic_sort_n<PAY, 8>(k, p);
  else if (size > 2) bitonic_sort_n<PAY, 4>(k, p);
  else if (size > 1) bitonic_sort_n<PAY, 2>(k, p);
}

template <bool PAY, class KT> __device__ __forceinline__ void bitonic_sort64(KT &k, uint32_t &p) {
  bitonic_sort_n<PAY, WAVE>(k, p);
}

// sorts a bitonic sequence held in each group of `size` lanes
template <bool PAY, class KT> __device__ __forceinline__ void bitonic_clean_groups(KT &k, uint32_t &p, int size) {
  if (size > 32) bitonic_clean_n<PAY, 64>(k, p);
  else if (size > 16) bitonic_clean_n<PAY, 32>(k, p);
  else if (size > 8) bitonic_clean_n<PAY, 16>(k, p);
  else if (size > 4) bitonic_clean_n<PAY, 8>(k, p);
  else if (size > 2) bitonic_clean_n<PAY, 4>(k, p);
  else if (size > 1) bitonic_clean_n<PAY, 2>(k, p);
}

template <bool PAY, class KT> __device__ __forceinline__ void bitonic_clean64(KT &k, uint32_t &p) {
  bitonic_clean_n<PAY, WAVE>(k, p);
}

// the smallest power of two >= v, for 2 <= v <= 2^30 (a count of leading zeros, not a loop)
__device__ __forceinline__ int pow2_at_least(int v) { return 1 << (32 - __clz(v - 1)); }

// Merge the staged hits into the best list.  Only the K best are ever read, so the list spans
// W = pow2 >= K lanes: the staged batch is sorted in groups of pow2 >= nb lanes, its W smallest are
// reversed onto lanes [0, W) and min'ed with the list (a bitonic sequence), and W lanes are cleaned.
template <bool PAY, class KT> __device__ __forceinline__ void topk_flush(TopKT<KT> &t, WaveLds &L) {
  static_assert(!PAY, "staged hits carry keys only");
  if (t.nb == 0) return;
  const int l = lane();
  wave_lds_sync();
  KT nk = l < t.nb ? (KT)L.bk[l] : ~(KT)0;
  uint32_t np = 0u;
  wave_lds_sync();
  const int sb = pow2_at_least(max(t.nb, 2));
  bitonic_sort_groups<PAY>(nk, np, sb);
  const int W = pow2_at_least(max(t.K, 2));
  if (t.cnt == 0) {
    // lanes [sb, 64) hold ~0 after the group sort; lanes >= W are never read
    t.key = nk;
    t.pay = np;
  } else {
    // lane l < W takes new[W-1-l] = new[l ^ (W-1)] (the batch's W smallest, reversed)
    const int ra = xor_addr(W - 1);
    KT rk = bperm(nk, ra);
    uint32_t rp = PAY ? bperm(np, ra) : 0u;
    if (l >= W) rk = ~(KT)0;
    if (rk < t.key) {
      t.key = rk;
      t.pay = rp;
    }
    bitonic_clean_groups<PAY>(t.key, t.pay, W);
    if (l >= W) t.key = ~(KT)0;
  }
  t.cnt = min(W, t.cnt + t.nb);
  t.nb = 0;
  t.kth = (t.cnt >= t.K) ? read_lane(t.key, t.K - 1) : t.bound;
}

template <bool PAY, class KT>
__device__ __forceinline__ void topk_offer(TopKT<KT> &t, WaveLds &L, bool hit, KT key, uint32_t) {
  static_assert(!PAY, "staged hits carry keys only");
  t.hits += (uint32_t)popc(ballot(hit));
  bool pass = hit && key < t.kth;
  uint64_t m = ballot(pass);
  int np = popc(m);
  if (np == 0) return;
  if (t.nb + np > WAVE) {
    topk_flush<PAY>(t, L);
    pass = hit && key < t.kth;
    m = ballot(pass);
    np = popc(m);
    if (np == 0) return;
  }
  if (pass) {
    const int pos = t.nb + popc(m & lanemask_lt());
    L.bk[pos] = key;
  }
  t.nb += np;
}

// Walk the cells of box [lo, hi] (minus box [ilo, ihi] when `inner`), calling sink(valid, hit, key, pos)
// uniformly on every lane for each group of 64 candidates.  radius_mode: hit = d2 <= r2; otherwise
// hit = d2 is not NaN (k-NN).
template <class Sink>
__device__ __forceinline__ void scan_box(const GridView &g, float qx, float qy, float qz, const int lo[3],
                                         const int hi[3], bool inner, const int ilo[3], const int ihi[3],
                                         float r2, bool radius_mode, WaveLds &L, Sink &sink) {
  const int l = lane();
  const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  if (ny <= 0 || nz <= 0 || hi[0] < lo[0]) return;
  const int nrows = ny * nz;
  const float rny = __frcp_rn((float)ny);
  const bool fq = nrows < (1 << 22);
  for (int r0 = 0; r0 < nrows; r0 += WAVE) {
    const int row = r0 + l;
    uint32_t s0 = 0, l0 = 0, s1 = 0, l1 = 0;
    if (row < nrows) {
      // (row + 0.5) / ny is at least 0.5 / ny from an integer, more than the rounding of the product for
      // row < 2^22, so the truncation is the exact quotient; larger boxes divide
      const int qz = fq ? (int)__fmul_rn((float)row + 0.5f, rny) : row / ny;
      const int iy = lo[1] + (row - qz * ny), iz = lo[2] + qz;
      // the dense table has fewer than 2^32 cells (Grid::build bounds it)
      const uint32_t base = ((uint32_t)iz * (uint32_t)g.gy + (uint32_t)iy) * (uint32_t)g.gx;
      if (inner && iy >= ilo[1] && iy <= ihi[1] && iz >= ilo[2] && iz <= ihi[2]) {
        if (ilo[0] > lo[0]) {
          s0 = g.cell_start[base + lo[0]];
          l0 = g.cell_start[base + ilo[0]] - s0;
        }
        if (ihi[0] < hi[0]) {
          s1 = g.cell_start[base + ihi[0] + 1];
          l1 = g.cell_start[base + hi[0] + 1] - s1;
        }
      } else {
        s0 = g.cell_start[base + lo[0]];
        l0 = g.cell_start[base + hi[0] + 1] - s0;
      }
    }
    const uint32_t len = l0 + l1;
    const uint32_t inc = wave_incl_scan(len);
    const uint32_t total = readlaneu(inc, WAVE - 1);
    if (total == 0) continue;
    wave_lds_sync();
    L.excl[l] = inc - len;
    L.s0[l] = s0;
    L.l0[l] = l0;
    L.s1[l] = s1;
    wave_lds_sync();
    // rows of this pass: the search below needs only as many halvings as its row count
    const int pass_rows = min(WAVE, nrows - r0);
    const int top = pass_rows > 16 ? WAVE / 2 : pass_rows > 4 ? 8 : 2;
    // candidate v of the pass -> its cell-order position
    auto position = [&](uint32_t v) {
      int o = 0;
#pragma unroll
      for (int st = WAVE / 2; st > 0; st >>= 1)
        if (st <= top && L.excl[o + st] <= v) o += st;
      const uint32_t off = v - L.excl[o];
      return off < L.l0[o] ? L.s0[o] + off : L.s1[o] + (off - L.l0[o]);
    };
    // two groups of 64 per step: the second group's loads are in flight while the first is offered
    for (uint32_t c0 = 0; c0 < total; c0 += 2 * WAVE) {
      const uint32_t v0 = c0 + (uint32_t)l, v1 = v0 + WAVE;
      const bool two = c0 + WAVE < total;
      const bool ok0 = v0 < total, ok1 = v1 < total;
      float x0 = 0.f, y0 = 0.f, z0 = 0.f, x1 = 0.f, y1 = 0.f, z1 = 0.f;
      uint32_t i0 = 0, i1 = 0, p0 = 0, p1 = 0;
      if (ok0) {
        p0 = position(v0);
        x0 = g.xs[p0];
        y0 = g.ys[p0];
        z0 = g.zs[p0];
        i0 = g.perm[p0];
      }
      if (two && ok1) {
        p1 = position(v1);
        x1 = g.xs[p1];
        y1 = g.ys[p1];
        z1 = g.zs[p1];
        i1 = g.perm[p1];
      }
      {
        float d2 = 0.0f;
        bool hit = false;
        if (ok0) {
          d2 = dist2(x0, y0, z0, qx, qy, qz);
          hit = radius_mode ? (d2 <= r2) : (d2 == d2);
        }
        sink(ok0, hit, pack_key(d2, i0), p0);
      }
      if (two) {
        float d2 = 0.0f;
        bool hit = false;
        if (ok1) {
          d2 = dist2(x1, y1, z1, qx, qy, qz);
          hit = radius_mode ? (d2 <= r2) : (d2 == d2);
        }
        sink(ok1, hit, pack_key(d2, i1), p1);
      }
    }
    wave_lds_sync();
  }
}

// One query of a wave's run of queries.  Lane j of the wave loads query j of the run and computes its window
// (the per-query set-up, divisions included, done once for the run, lane-parallel); of_lane(j) hands query j to
// the whole wave (read from lane j: wave-uniform values).
struct LaneQuery {
  float x = 0.f, y = 0.f, z = 0.f;
  uint32_t row = 0;
  int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
  __device__ __forceinline__ void load(const GridView &g, float qx, float qy, float qz, uint32_t out_row, float r) {
    x = qx;
    y = qy;
    z = qz;
    row = out_row;
    window3(g, x, y, z, r, lo, hi);
  }
  __device__ __forceinline__ LaneQuery of_lane(int j) const {
    LaneQuery q;
    q.x = readlanef(x, j);
    q.y = readlanef(y, j);
    q.z = readlanef(z, j);
    q.row = readlaneu(row, j);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      q.lo[a] = __builtin_amdgcn_readlane(lo[a], j);
      q.hi[a] = __builtin_amdgcn_readlane(hi[a], j);
    }
    return q;
  }
};

// queries per wave for a run-of-queries kernel: up to 64, fewer when the count would leave the chip
// (256 CUs x 32 wave slots) less than two rounds of waves; the "queries_per_wave" knob overrides it
inline int queries_per_wave(size_t nq) {
  const int forced = tune_queries_per_wave();
  if (forced > 0) return forced;
  const size_t target_waves = 16384;
  return (int)std::max<size_t>(1, std::min<size_t>(WAVE, nq / target_waves));
}

// consecutive queries per wave for the k-NN means: a run ends the chip's work unevenly (the last runs of
// the last round), so runs stay short until there are dozens of rounds of resident waves: 2 up to a quarter
// million queries, nq / 2^17 above, at most 16; a positive `forced` (the queries_per_wave knob) overrides it
__host__ __device__ inline int knn_run_length(size_t nq, int forced) {
  if (forced > 0) return forced;
  const size_t r = nq >> 17;
  return (int)(r < 2 ? 2 : r > 16 ? 16 : r);
}

// A wave's copy of the candidates of one window: x, y, z and the original index (as float bits) of every point
// of the box, in slice order.  Consecutive queries of a run are mostly in one cell and then have the same
// window, so one staging serves them all and each query reads its candidates from LDS, 64 per instruction.
// CAP points at most: 256 (4 KB per wave), or 224 for kernels that fit eight waves per SIMD only with at most
// 20 KB of LDS per block of four waves.
template <int CAP> struct StageLdsT {
  static constexpr int CAPACITY = CAP;
  static constexpr int GROUPS = (CAP + WAVE - 1) / WAVE;  // groups of 64 staged points
  float4 c[CAP];
};
using StageLds = StageLdsT<256>;
using StageLdsSmall = StageLdsT<224>;

// Stage the points of box [lo, hi] into S; returns their count, or -1 when the box has more than 64 (y, z)
// rows or more points than S holds (the caller then walks the box with scan_box).
template <class ST>
__device__ __forceinline__ int stage_box(const GridView &g, const int lo[3], const int hi[3], WaveLds &L, ST &S) {
  const int l = lane();
  const int ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
  if (ny <= 0 || nz <= 0 || hi[0] < lo[0]) return 0;
  const int nrows = ny * nz;
  if (nrows > WAVE) return -1;
  uint32_t s0 = 0, l0 = 0;
  if (l < nrows) {
    const int qz = (int)__fmul_rn((float)l + 0.5f, __frcp_rn((float)ny));  // exact quotient (see scan_box)
    const int iy = lo[1] + (l - qz * ny), iz = lo[2] + qz;
    const uint32_t base = ((uint32_t)iz * (uint32_t)g.gy + (uint32_t)iy) * (uint32_t)g.gx;
    s0 = g.cell_start[base + lo[0]];
    l0 = g.cell_start[base + hi[0] + 1] - s0;
  }
  const uint32_t inc = wave_incl_scan(l0);
  const uint32_t total = readlaneu(inc, WAVE - 1);
  if (total > (uint32_t)ST::CAPACITY) return -1;
  wave_lds_sync();
  L.excl[l] = inc - l0;  // rows past nrows hold `total`, so the search below never selects them
  L.s0[l] = s0;
  wave_lds_sync();
  const int top = nrows > 16 ? WAVE / 2 : nrows > 4 ? 8 : 2;
  // every group's gathers are issued before the first is stored, so the staging waits for one round trip to
  // memory instead of one per group
  constexpr int NG = ST::GROUPS;
  float4 c[NG];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    const uint32_t v = (uint32_t)(gi * WAVE + l);
    c[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((uint32_t)(gi * WAVE) < total && v < total) {
      int o = 0;
#pragma unroll
      for (int st = WAVE / 2; st > 0; st >>= 1)
        if (st <= top && L.excl[o + st] <= v) o += st;
      const uint32_t pos = L.s0[o] + (v - L.excl[o]);
      c[gi] = make_float4(g.xs[pos], g.ys[pos], g.zs[pos], __uint_as_float(g.perm[pos]));
    }
  }
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    const uint32_t v = (uint32_t)(gi * WAVE + l);
    if (v < total) S.c[v] = c[gi];
  }
  wave_lds_sync();
  return (int)total;
}

// The staged counterpart of scan_box in radius mode: sink(valid, hit, key, 0) for each group of 64 candidates.
template <class Sink, class ST>
__device__ __forceinline__ void scan_staged(const ST &S, int total, float qx, float qy, float qz, float r2,
                                            Sink &sink) {
  for (int c0 = 0; c0 < total; c0 += WAVE) {
    const int v = c0 + lane();
    const bool valid = v < total;
    float d2 = 0.0f;
    uint32_t idx = 0;
    bool hit = false;
    if (valid) {
      const float4 p = S.c[v];
      d2 = dist2(p.x, p.y, p.z, qx, qy, qz);
      idx = __float_as_uint(p.w);
      hit = d2 <= r2;
    }
    sink(valid, hit, pack_key(d2, idx), 0u);
  }
}

// The radius hits of a staged window for callers that read the neighbour set and not its order (covariances):
// when at most K staged points lie within r2 of the query they ARE the list capped at K, so no selection or sort
// is needed.  With more hits, the K nearest are those at or below a d2 threshold t with exactly K hits at or
// below it, found by interpolating the counts of the hits' d2 (every other step bisecting the bit patterns).
// Returns the neighbour count (min(H, K)), lane j holding neighbour j's coordinates (staging order); -1 when no
// threshold separates exactly K (the K-th and (K+1)-th distances are equal: the caller then orders the list by
// (d2, idx)).  Uses L.excl as scratch.
template <class ST>
__device__ __forceinline__ int staged_hits_upto(const ST &S, int total, float qx, float qy, float qz, float r2,
                                                int K, WaveLds &L, float &hx, float &hy, float &hz) {
  constexpr int NG = ST::GROUPS;
  const int l = lane();
  uint64_t m[NG];
  uint32_t d[NG];  // hits' d2 bits, all-ones for the rest
  int H = 0;
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    m[gi] = 0;
    d[gi] = ~0u;
    if (gi * WAVE < total) {
      const int v = gi * WAVE + l;
      bool hit = false;
      if (v < total) {
        const float4 p = S.c[v];
        const float d2 = dist2(p.x, p.y, p.z, qx, qy, qz);
        hit = d2 <= r2;
        if (hit) d[gi] = __float_as_uint(d2);
      }
      m[gi] = ballot(hit);
      H += popc(m[gi]);
    }
  }
  if (H > K) {
    auto count = [&](uint32_t t) {
      int c = 0;
#pragma unroll
      for (int gi = 0; gi < NG; ++gi)
        if (gi * WAVE < total) c += popc(ballot(d[gi] <= t));
      return c;
    };
    // d2 >= +0, so the bit patterns order as the values; c(lo) < K < c(hi) throughout
    uint32_t lo = 0, hi = __float_as_uint(r2), tau = 0;
    int clo = count(0), chi = H;
    bool found = clo == K;
    if (clo < K) {
      for (int it = 0; it < 64 && hi - lo > 1; ++it) {
        uint32_t t = lo + (hi - lo) / 2;
        if ((it & 1) == 0) {
          const float flo = __uint_as_float(lo), fhi = __uint_as_float(hi);
          const float f = flo + (fhi - flo) * (((float)K - 0.5f - (float)clo) / (float)(chi - clo));
          const uint32_t ti = __float_as_uint(f);
          if (f == f && ti > lo && ti < hi) t = ti;
        }
        const int ct = count(t);
        if (ct == K) {
          tau = t;
          found = true;
          break;
        }
        if (ct < K) {
          lo = t;
          clo = ct;
        } else {
          hi = t;
          chi = ct;
        }
      }
    }
    if (!found) return -1;
#pragma unroll
    for (int gi = 0; gi < NG; ++gi)
      if (gi * WAVE < total) m[gi] = ballot(d[gi] <= tau);
    H = K;
  }
  int base = 0;
  wave_lds_sync();
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    if ((m[gi] >> l) & 1ull) L.excl[base + popc(m[gi] & lanemask_lt())] = (uint32_t)(gi * WAVE + l);
    base += popc(m[gi]);
  }
  wave_lds_sync();
  hx = hy = hz = 0.0f;
  if (l < H) {
    const float4 p = S.c[L.excl[l]];
    hx = p.x;
    hy = p.y;
    hz = p.z;
  }
  wave_lds_sync();
  return H;
}

// Window cache of a run of queries: restages only when a query's window differs from the staged one.
// `tag` names the grid the window belongs to (the cloud of a batch; 0 for a single grid).
struct StagedWindow {
  int lo[3] = {1, 1, 1}, hi[3] = {0, 0, 0};
  int tag = -1;
  int total = -2;  // -2: nothing staged; -1: window too large to stage
  template <class ST>
  __device__ __forceinline__ void ensure(const GridView &g, const int qlo[3], const int qhi[3], int qtag,
                                         WaveLds &L, ST &S) {
    if (total != -2 && qtag == tag && qlo[0] == lo[0] && qlo[1] == lo[1] && qlo[2] == lo[2] &&
        qhi[0] == hi[0] && qhi[1] == hi[1] && qhi[2] == hi[2])
      return;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = qlo[a];
      hi[a] = qhi[a];
    }
    tag = qtag;
    total = stage_box(g, lo, hi, L, S);
  }
};

template <bool PAY, class KT = uint64_t> struct TopKSink {
  TopKT<KT> &t;
  WaveLds &L;
  __device__ __forceinline__ void operator()(bool, bool hit, uint64_t key, uint32_t pos) {
    topk_offer<PAY>(t, L, hit, as_key<KT>(key), pos);
  }
};

// Radius whose window3 holds every point p with fl(d2(p, q)) <= b: the hardware square root (within 2 ulp)
// widened by 2^-20 and one ulp, so fl(R * R) >= b.
__device__ __forceinline__ float cert_radius(float b) {
  return next_up(__fmul_rn(__builtin_amdgcn_sqrtf(b), 1.0f + 0x1p-20f));
}

// the smallest key above every key whose d2 is at most b (b finite, >= 0)
template <class KT> __device__ __forceinline__ KT bound_key(float b) {
  return as_key<KT>((uint64_t)(__float_as_uint(b) + 1u) << 32);
}

// An upper bound on the K-th smallest d2 at query q from the previous query p of the same cloud, whose K-th
// smallest d2 is pd2: p's K nearest lie within sqrt(pd2) + |q - p| of q (triangle inequality).  Computed
// d2 values are within a relative 5 * 2^-24 of the true squared distances, and the evaluation below loses
// at most 12 * 2^-24 more (two hardware square roots among them), so the 2^-16 margin covers both.  +inf (no bound) when the step exceeds half the
// previous K-th distance (the bound would admit too many candidates to pay), for tiny or non-finite pd2.
__device__ __forceinline__ float knn_bound_from(float qx, float qy, float qz, float px, float py, float pz,
                                                float pd2) {
  if (!(pd2 >= 0x1p-60f && pd2 < 0x1p100f)) return __builtin_inff();
  const float dx = __fsub_rn(qx, px), dy = __fsub_rn(qy, py), dz = __fsub_rn(qz, pz);
  const float dd =
      __builtin_amdgcn_sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz)));
  const float dp = __builtin_amdgcn_sqrtf(pd2);
  if (!(dd <= 0.5f * dp)) return __builtin_inff();
  const float sm = __fadd_rn(dp, dd);
  return __fmul_rn(__fmul_rn(sm, sm), 1.0f + 0x1p-16f);
}

// Exact k nearest by (d2, idx): the cells within h0 of the query's cell (3x3x3 for h0 = 1), doubling growth
// until k candidates are known, then the box certified by window3(q, R) with fl(R*R) > d2_k: every point
// outside it has a larger d2, so the k best are final.  With a finite bound_d2 (at least the K-th smallest d2), the first
// box is the one certified for bound_d2 and only hits within it are staged: once it yields K entries the
// certification holds at once; if it yields fewer, the search starts over without the bound.
template <bool PAY, class KT>
__device__ __forceinline__ void knn_search(const GridView &g, float qx, float qy, float qz, TopKT<KT> &t,
                                           WaveLds &L, float bound_d2 = __builtin_inff(), int h0 = 1) {
  const int gd[3] = {g.gx, g.gy, g.gz};
  int lo[3], hi[3];
  TopKSink<PAY, KT> sink{t, L};
  const bool bounded = bound_d2 < FLT_MAX;
  if (bounded) {
    const float R = cert_radius(bound_d2);
    window3(g, qx, qy, qz, R, lo, hi);
    t.bound = t.kth = bound_key<KT>(bound_d2);
  } else {
    // the query's cell up to rounding: any box is a valid start, the certification below makes the result
    // exact
    const int c[3] = {floor_clamped(cell_coord(qx, g.mnx, g.icell), g.gx),
                      floor_clamped(cell_coord(qy, g.mny, g.icell), g.gy),
                      floor_clamped(cell_coord(qz, g.mnz, g.icell), g.gz)};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = max(c[a] - h0, 0);
      hi[a] = min(c[a] + h0, gd[a] - 1);
    }
  }
  scan_box(g, qx, qy, qz, lo, hi, false, lo, hi, 0.0f, false, L, sink);
  // a bounded box with fewer than K entries (a bound below the K-th d2): forget it and grow from the box,
  // rescanning it whole
  bool rescan = false;
  if (bounded) {
    topk_flush<PAY>(t, L);
    if (t.cnt < t.K) {
      t.init(t.K);
      rescan = true;
    }
  }
  int step = 1;
  for (;;) {
    const bool whole = lo[0] == 0 && lo[1] == 0 && lo[2] == 0 && hi[0] == gd[0] - 1 && hi[1] == gd[1] - 1 &&
                       hi[2] == gd[2] - 1;
    // while fewer than K candidates are known the box only grows: merging can wait
    if (whole || t.cnt + t.nb >= t.K) topk_flush<PAY>(t, L);
    if (whole && !rescan) break;
    int nlo[3], nhi[3];
    if (t.cnt >= t.K) {
      const float dk = kd2(t.kth);
      const float R = cert_radius(dk);
      if (!(R < FLT_MAX)) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          nlo[a] = 0;
          nhi[a] = gd[a] - 1;
        }
      } else {
        int wlo[3], whi[3];
        window3(g, qx, qy, qz, R, wlo, whi);
        if (wlo[0] >= lo[0] && wlo[1] >= lo[1] && wlo[2] >= lo[2] && whi[0] <= hi[0] && whi[1] <= hi[1] &&
            whi[2] <= hi[2])
          break;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          nlo[a] = min(lo[a], wlo[a]);
          nhi[a] = max(hi[a], whi[a]);
        }
      }
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        nlo[a] = max(lo[a] - step, 0);
        nhi[a] = min(hi[a] + step, gd[a] - 1);
      }
      step = step < (1 << 20) ? step * 2 : step;
    }
    scan_box(g, qx, qy, qz, nlo, nhi, !rescan, lo, hi, 0.0f, false, L, sink);
    rescan = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = nlo[a];
      hi[a] = nhi[a];
    }
  }
}

// The SOR means' sequential sums, deferred: each query's sorted distances (entry j on lane j) go to a row of
// LDS, and once SUM_SLOTS rows are queued lane i sums row i, entries 1..k in order from 0.  The k dependent
// adds then serve SUM_SLOTS queries instead of one (and need no lane-to-scalar reads); the sums are the same
// sequence of fp32 adds, so the means are bit for bit those of a per-query sum.  Rows are WAVE + 1 words
// apart so that the lanes of one read hit distinct banks.
constexpr int SUM_SLOTS = 4;
struct SumLds {
  float v[SUM_SLOTS][WAVE + 1];
  uint32_t dst[SUM_SLOTS];
};

struct DeferredMeans {
  int n = 0;  // rows queued (wave-uniform)
  // queue one query: `sq` is lane j's entry j (lanes beyond k are never read), `dst` its output index
  __device__ __forceinline__ void add(SumLds &S, float sq, uint32_t dst, int k, float *__restrict__ means) {
    S.v[n][lane()] = sq;
    if (lane() == 0) S.dst[n] = dst;
    if (++n == SUM_SLOTS) flush(S, k, means);
  }
  __device__ __forceinline__ void flush(SumLds &S, int k, float *__restrict__ means) {
    if (n == 0) return;
    wave_lds_sync();
    const int l = lane();
    if (l < n) {
      float s = 0.0f;
      for (int j = 1; j <= k; ++j) s = __fadd_rn(s, S.v[l][j]);
      means[S.dst[l]] = __fdiv_rn(s, __int2float_rn(k));
    }
    wave_lds_sync();
    n = 0;
  }
};

// Several fp64 sums over the wave at once, reduce-scatter style: at each butterfly distance the lanes split
// the values between the two halves of each pair (each keeps one and adds its partner's copy), so a value is
// combined on half as many lanes as a separate butterfly would use; the last steps are a full butterfly within
// the lane group that holds the value.  fp64 adds issue at the rate of unpacked fp32 ones on this GPU, so the
// sums cost a third of compensated fp32 sums and are more accurate.  Both lanes of a pair add the same two
// numbers, so every lane of a group ends with the same result.
__device__ __forceinline__ double readlaned(double v, int l) {
  return __longlong_as_double((long long)readlane64((uint64_t)__double_as_longlong(v), l));
}

// lanes with `upper` clear keep value lo, the others value hi; each adds the partner's copy of what it keeps
__device__ __forceinline__ double rs_step64(double lo, double hi, bool upper, int addr) {
  const double give = upper ? lo : hi;  // the value the partner keeps
  return (upper ? hi : lo) + bperm(give, addr);
}

// out[i] = the sum over the wave of in[i], i < 3 (wave-uniform results)
__device__ __forceinline__ void wave_sum3_f64(const double in[3], double out[3]) {
  const int l = lane();
  const int a32 = xor_addr(32), a16 = xor_addr(16);
  const double w0 = rs_step64(in[0], in[2], (l & 32) != 0, a32);  // lanes < 32: value 0; >= 32: value 2
  double w1 = in[1] + bperm(in[1], a32);                          // lanes < 32: value 1
  if (l & 32) w1 = w0;
  double y = rs_step64(w0, w1, (l & 16) != 0, a16);  // lanes 0-15: value 0, 16-31: value 1, 32-63: value 2
#pragma unroll
  for (int d = 8; d > 0; d >>= 1) y += bperm(y, xor_addr(d));
  out[0] = readlaned(y, 0);
  out[1] = readlaned(y, 16);
  out[2] = readlaned(y, 32);
}

// out[i] = the sum over the wave of in[i], i < 6 (wave-uniform results)
__device__ __forceinline__ void wave_sum6_f64(const double in[6], double out[6]) {
  const int l = lane();
  const int a32 = xor_addr(32), a16 = xor_addr(16), a8 = xor_addr(8);
  double w[3];  // lanes < 32: values 0, 1, 2; lanes >= 32: values 3, 4, 5
#pragma unroll
  for (int i = 0; i < 3; ++i) w[i] = rs_step64(in[i], in[i + 3], (l & 32) != 0, a32);
  const double x0 = rs_step64(w[0], w[2], (l & 16) != 0, a16);  // bit 4 clear: value 0 (3); set: value 2 (5)
  double x1 = w[1] + bperm(w[1], a16);                          // bit 4 clear: value 1 (4)
  if (l & 16) x1 = x0;
  // lanes 0-7 value 0, 8-15 value 1, 16-31 value 2, 32-39 value 3, 40-47 value 4, 48-63 value 5
  double y = rs_step64(x0, x1, (l & 8) != 0, a8);
#pragma unroll
  for (int d = 4; d > 0; d >>= 1) y += bperm(y, xor_addr(d));
  out[0] = readlaned(y, 0);
  out[1] = readlaned(y, 8);
  out[2] = readlaned(y, 16);
  out[3] = readlaned(y, 32);
  out[4] = readlaned(y, 40);
  out[5] = readlaned(y, 48);
}

// CSR fill of radius lists (rows = original query index; qx == nullptr: the grid's own points), each row
// sorted by (d2, idx); pos (may be null) receives the cell-order position of each neighbour.
int radius_csr(const Grid &G, const float *qx, const float *qy, const float *qz, size_t nq, float r,
               const int64_t *offsets, uint64_t *keys, uint32_t *pos, hipStream_t s);

}  // namespace rvp
