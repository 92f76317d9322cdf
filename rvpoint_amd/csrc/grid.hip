// grid.hip -- SpatialHash build on the GPU: bbox, cell keys, stable radix sort, cell-ordered SoA, cell table.
//
// Reference: SpatialHash::setInputCloud / build (src/search/spatial_hashing.cpp:76-163).  The reference
// buckets point indices in an unordered_map keyed by ix + iy*73856093 + iz*19349663; here the cells of a
// dense grid are sorted by a linear id so that every run of x-adjacent cells is one contiguous slice of
// the cell-ordered point arrays.  Query results do not depend on the cell layout (see query.hip).
#include <algorithm>
#include <vector>

#include "grid.hpp"
#include "prims.hpp"

namespace rvp {

// ------------------------------------------------------------------------------------------------
// bbox
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bbox_partial_k(const float *__restrict__ x, const float *__restrict__ y,
                                                       const float *__restrict__ z, size_t n,
                                                       float *__restrict__ part) {
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float a = x[i], b = y[i], c = z[i];
    // strict comparisons: NaN never updates, as the reference's scalar loop (spatial_hashing.cpp:131-138)
    if (a < v[0]) v[0] = a;
    if (b < v[1]) v[1] = b;
    if (c < v[2]) v[2] = c;
    if (a > v[3]) v[3] = a;
    if (b > v[4]) v[4] = b;
    if (c > v[5]) v[5] = c;
  }
  __shared__ float sm[4][6];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float o = shfl_xor(v[j], d);
      if (o < v[j]) v[j] = o;
      float p = shfl_xor(v[j + 3], d);
      if (p > v[j + 3]) v[j + 3] = p;
    }
  }
  if (lane() == 0)
    for (int j = 0; j < 6; ++j) sm[threadIdx.x >> 6][j] = v[j];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w)
      for (int j = 0; j < 3; ++j) {
        if (sm[w][j] < v[j]) v[j] = sm[w][j];
        if (sm[w][j + 3] > v[j + 3]) v[j + 3] = sm[w][j + 3];
      }
    for (int j = 0; j < 6; ++j) part[blockIdx.x * 6 + j] = v[j];
  }
}

// one wave: lanes stride over the block partials, then a butterfly (min / max are exact in any order; the
// partials hold no NaN)
__global__ __launch_bounds__(64) void bbox_final_k(const float *__restrict__ part, int nb, float *__restrict__ out) {
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int b = threadIdx.x; b < nb; b += WAVE)
    for (int j = 0; j < 3; ++j) {
      v[j] = fminf(v[j], part[b * 6 + j]);
      v[j + 3] = fmaxf(v[j + 3], part[b * 6 + j + 3]);
    }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1)
    for (int j = 0; j < 3; ++j) {
      v[j] = fminf(v[j], shfl_xor(v[j], d));
      v[j + 3] = fmaxf(v[j + 3], shfl_xor(v[j + 3], d));
    }
  if (threadIdx.x == 0)
    for (int j = 0; j < 6; ++j) out[j] = v[j];
}

int bbox_minmax(const float *x, const float *y, const float *z, size_t n, float *out6, hipStream_t s) {
  const int nb = (int)std::min<size_t>(1024, std::max<size_t>(1, ceil_div(n, 256 * 8)));
  Tmp part;
  RVP_TRY(part.alloc(nb * 6 * sizeof(float), s));
  RVP_LAUNCH("bbox_partial_k", s, bbox_partial_k<<<nb, 256, 0, s>>>(x, y, z, n, part.as<float>()));
  RVP_LAUNCH("bbox_final_k", s, bbox_final_k<<<1, 64, 0, s>>>(part.as<float>(), nb, out6));
  return RVP_OK;
}

// ------------------------------------------------------------------------------------------------
// grid geometry (host): the reference's bbox margin and the cells per axis for an edge
// ------------------------------------------------------------------------------------------------
// spatial_hashing.cpp:141-146: min -= cell * 0.01f, max += cell * 0.01f (fp32, as on the host there)
static void margin_bbox(const float mm[6], float cell, float bb[6]) {
  const float m = cell * 0.01f;
  for (int j = 0; j < 3; ++j) {
    bb[j] = mm[j] - m;
    bb[j + 3] = mm[j + 3] + m;
  }
}

// cells per axis = floor((data_max - origin) / edge) + 1, the same fp32 expression the device cell
// function evaluates; points beyond the last cell would be clamped into it, never dropped
static double cells_on_axis(float data_max, float origin, float edge) {
  const float t = (data_max - origin) / edge;
  return std::floor((double)t) + 1.0;
}

// one byte per cell; racing plain stores of the same value need no atomics (read first: most cells are
// already marked)
__global__ void occupancy_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                            size_t n, float mnx, float mny, float mnz, float c, int gx, int gy, int gz,
                            uint8_t *__restrict__ flags) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = (uint32_t)cell_clamped(x[i], mnx, c, gx) +
                     (uint32_t)gx * ((uint32_t)cell_clamped(y[i], mny, c, gy) +
                                     (uint32_t)gy * (uint32_t)cell_clamped(z[i], mnz, c, gz));
  if (!flags[k]) flags[k] = 1;
}

// nonzero bytes of a word
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t w) {
  w |= w >> 1;
  w |= w >> 2;
  w |= w >> 4;
  return __popc(w & 0x01010101u);
}

// marked cells (flags padded to a multiple of 4 bytes with zeros)
__global__ __launch_bounds__(256) void count_marked_k(const uint32_t *__restrict__ words, size_t nw,
                                                       uint32_t *__restrict__ count) {
  uint32_t c = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (size_t)gridDim.x * blockDim.x)
    c += nonzero_bytes(words[i]);
  c = wave_sum(c);
  if (lane() == 0 && c) atomicAdd(count, c);
}

// also each point as one 16-byte record in input order, so that the gather after the sort reads one sector
// per point instead of three
__global__ void keys_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                       size_t n, float mnx, float mny, float mnz, float c, int gx, int gy, int gz,
                       uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, float4 *__restrict__ xyz) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a = x[i], b = y[i], d = z[i];
  int ix = cell_clamped(a, mnx, c, gx);
  int iy = cell_clamped(b, mny, c, gy);
  int iz = cell_clamped(d, mnz, c, gz);
  keys[i] = (uint32_t)ix + (uint32_t)gx * ((uint32_t)iy + (uint32_t)gy * (uint32_t)iz);
  vals[i] = (uint32_t)i;
  xyz[i] = make_float4(a, b, d, 0.0f);
}

// cell-order SoA, and the inverse permutation rank[perm[i]] = i
__global__ void gather_k(const float4 *__restrict__ xyz, const uint32_t *__restrict__ perm, size_t n,
                         float *__restrict__ xs, float *__restrict__ ys, float *__restrict__ zs,
                         uint32_t *__restrict__ rank) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t p = perm[i];
  const float4 v = xyz[p];
  xs[i] = v.x;
  ys[i] = v.y;
  zs[i] = v.z;
  rank[p] = (uint32_t)i;
}

// cs[k] = length of the run of key k, written by the run's first position (the table is zeroed before): the
// exclusive scan of these counts is the cell-start table.  The run end comes from an exponential, then binary
// search over the sorted keys.
__global__ void run_lengths_k(const uint32_t *__restrict__ keys, size_t n, uint32_t *__restrict__ cs) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (i > 0 && keys[i - 1] == k) return;
  size_t lo = i, hi, step = 1;  // keys[lo] == k
  for (;;) {
    hi = lo + step;
    if (hi >= n || keys[hi] != k) break;
    lo = hi;
    step <<= 1;
  }
  if (hi > n) hi = n;  // keys[hi] != k or hi == n
  while (hi - lo > 1) {
    const size_t m = lo + (hi - lo) / 2;
    if (keys[m] == k) lo = m;
    else hi = m;
  }
  cs[k] = (uint32_t)(hi - i);
}

// occupied cells: c with cs[c + 1] > cs[c]
__global__ __launch_bounds__(256) void count_occupied_k(const uint32_t *__restrict__ cs, uint32_t ncells,
                                                         uint32_t *__restrict__ out) {
  uint32_t cnt = 0;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x)
    cnt += cs[c + 1] > cs[c] ? 1u : 0u;
  cnt = wave_sum(cnt);
  __shared__ uint32_t sm[4];
  if (lane() == 0) sm[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, sm[0] + sm[1] + sm[2] + sm[3]);
}

// ------------------------------------------------------------------------------------------------
// Grid::build
// ------------------------------------------------------------------------------------------------
int Grid::build(const float *x_, const float *y_, const float *z_, size_t n_, float cell_, hipStream_t s) {
  RVP_CHECK(cell_ > 0.0f && cell_ < FLT_MAX, RVP_E_ARG, "cell size must be positive and finite");
  RVP_CHECK(n_ < 0xfffffff0ull, RVP_E_UNSUPPORTED, "more than 2^32-16 points");
  x = x_; y = y_; z = z_; n = n_; req_cell = cell_;
  built = false;
  stream = s;
  if (n == 0) {
    std::fill(bb, bb + 6, 0.f);
    cell = cell_;
    dims[0] = dims[1] = dims[2] = 1;
    ncells = 1;
    RVP_TRY(cstart.alloc(2 * sizeof(uint32_t), s, persistent));
    RVP_HIP(hipMemsetAsync(cstart.p, 0, 2 * sizeof(uint32_t), s));
    RVP_TRY(xs.alloc(4, s, persistent));
    RVP_TRY(ys.alloc(4, s, persistent));
    RVP_TRY(zs.alloc(4, s, persistent));
    RVP_TRY(perm.alloc(4, s, persistent));
    RVP_TRY(rank.alloc(4, s, persistent));
    built = true;
    return RVP_OK;
  }
  RVP_CHECK(x && y && z, RVP_E_ARG, "null coordinate pointer");
  float mm[6];
  RVP_TRY(minmax_host(mm, s));
  return build_from(mm, cell_, s);
}

int Grid::minmax_host(float mm[6], hipStream_t s) {
  Tmp d;
  RVP_TRY(d.alloc(6 * sizeof(float), s));
  RVP_TRY(bbox_minmax(x, y, z, n, d.as<float>(), s));
  RVP_HIP(hipMemcpyAsync(mm, d.p, 6 * sizeof(float), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  return RVP_OK;
}

int Grid::build_from(const float mm[6], float cell_, hipStream_t s) {
  req_cell = cell_;
  margin_bbox(mm, cell_, bb);
  // all-NaN cloud: min > max; every point then lands in cell 0 and never matches a query
  const bool degenerate = !(bb[0] <= bb[3] && bb[1] <= bb[4] && bb[2] <= bb[5]);
  // cell ids are 32-bit (sort keys, table offsets): the table stays below 2^32 cells
  const double limit = std::min<double>(std::max<double>(16777216.0, 4.0 * (double)n), 4294967040.0);
  cell = cell_;
  dims[0] = dims[1] = dims[2] = 1;
  if (!degenerate) {
    bool fits = false;
    for (int k = 0; k < 64 && cell < FLT_MAX; ++k) {
      const double d0 = cells_on_axis(mm[3], bb[0], cell), d1 = cells_on_axis(mm[4], bb[1], cell),
                   d2 = cells_on_axis(mm[5], bb[2], cell);
      if (d0 >= 1 && d1 >= 1 && d2 >= 1 && d0 <= 1e9 && d1 <= 1e9 && d2 <= 1e9 && d0 * d1 * d2 <= limit) {
        dims[0] = (int)d0;
        dims[1] = (int)d1;
        dims[2] = (int)d2;
        fits = true;
        break;
      }
      cell = cell * 2.0f;
    }
    RVP_CHECK(fits, RVP_E_UNSUPPORTED, "grid: coordinate extent too large for a dense cell table");
  }
  ncells = (uint64_t)dims[0] * dims[1] * dims[2];

  Tmp kt, vt;
  RVP_TRY(keys.alloc(n * sizeof(uint32_t), s, persistent));
  RVP_TRY(perm.alloc(n * sizeof(uint32_t), s, persistent));
  RVP_TRY(kt.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(vt.alloc(n * sizeof(uint32_t), s));
  const unsigned nb = ceil_div(n, 256);
  Tmp xyz;
  RVP_TRY(xyz.alloc(n * sizeof(float4), s));
  RVP_LAUNCH("keys_k", s, keys_k<<<nb, 256, 0, s>>>(x, y, z, n, bb[0], bb[1], bb[2], cell, dims[0], dims[1],
                                                    dims[2], keys.as<uint32_t>(), perm.as<uint32_t>(),
                                                    xyz.as<float4>()));
  uint32_t *kp = keys.as<uint32_t>(), *vp = perm.as<uint32_t>(), *ktp = kt.as<uint32_t>(), *vtp = vt.as<uint32_t>();
  RVP_TRY(radix_sort_pairs_u32(&kp, &vp, &ktp, &vtp, n, bits_for(ncells - 1), s, nullptr, nullptr, n >= (1u << 19)));
  if (kp != keys.as<uint32_t>()) {  // odd pass count: the sorted data sits in the temporaries
    RVP_HIP(hipMemcpyAsync(keys.p, kp, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    RVP_HIP(hipMemcpyAsync(perm.p, vp, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
  RVP_TRY(xs.alloc(n * sizeof(float), s, persistent));
  RVP_TRY(ys.alloc(n * sizeof(float), s, persistent));
  RVP_TRY(zs.alloc(n * sizeof(float), s, persistent));
  RVP_TRY(rank.alloc(n * sizeof(uint32_t), s, persistent));
  RVP_LAUNCH("gather_k", s, gather_k<<<nb, 256, 0, s>>>(xyz.as<float4>(), perm.as<uint32_t>(), n, xs.as<float>(),
                                                        ys.as<float>(), zs.as<float>(), rank.as<uint32_t>()));

  // cell-start table: run length of every occupied cell, then one exclusive scan over the table (cells past
  // the last key get n); per-run writes and a streaming scan instead of per-run gap filling
  RVP_TRY(cstart.alloc((ncells + 1) * sizeof(uint32_t), s, persistent));
  RVP_HIP(hipMemsetAsync(cstart.p, 0, (ncells + 1) * sizeof(uint32_t), s));
  RVP_LAUNCH("run_lengths_k", s, run_lengths_k<<<nb, 256, 0, s>>>(keys.as<uint32_t>(), n, cstart.as<uint32_t>()));
  RVP_TRY(exclusive_scan_u32(cstart.as<uint32_t>(), cstart.as<uint32_t>(), ncells + 1, nullptr, s));
  built = true;
  return RVP_OK;
}

size_t Grid::occupied_cells() const {
  if (!built || !cstart.p || n == 0) return 0;
  Tmp cnt;
  if (cnt.alloc(sizeof(uint32_t), stream) != RVP_OK) return 0;
  if (hipMemsetAsync(cnt.p, 0, sizeof(uint32_t), stream) != hipSuccess) return 0;
  const unsigned nb = (unsigned)std::min<size_t>(1024, std::max<size_t>(1, ceil_div(ncells, 256)));
  count_occupied_k<<<nb, 256, 0, stream>>>(cstart.as<uint32_t>(), (uint32_t)ncells, cnt.as<uint32_t>());
  uint32_t v = 0;
  if (hipMemcpyAsync(&v, cnt.p, sizeof(v), hipMemcpyDeviceToHost, stream) != hipSuccess) return 0;
  if (hipStreamSynchronize(stream) != hipSuccess) return 0;
  return v;
}

// Cell edge for k-nearest queries.  A first edge c0 from the bbox volume (k points per cell if the cloud
// filled it) gives an occupancy count; with m points per occupied cell the surface density is about
// m / c0^2, so the k-th neighbour distance is about R = c0 sqrt(k / (pi m)), and the edge is R / div
// (div = tuning knob knn_cell_div, default 2.4):
// the certified box then spans about 6 cells per axis, which keeps the candidates near 2-3x the k
// neighbours for surfaces and volumes alike.
int Grid::build_for_knn(const float *x_, const float *y_, const float *z_, size_t n_, int k, hipStream_t s) {
  x = x_; y = y_; z = z_; n = n_;
  if (n_ == 0) return build(x_, y_, z_, n_, 1.0f, s);
  stream = s;
  float mm[6];
  RVP_TRY(minmax_host(mm, s));
  double ext[3], vol = 1.0, emax = 0.0;
  for (int j = 0; j < 3; ++j) {
    ext[j] = (mm[j + 3] >= mm[j]) ? (double)mm[j + 3] - (double)mm[j] : 0.0;
    emax = std::max(emax, ext[j]);
  }
  if (!(emax > 0.0) || !(emax < 1e30)) return build_from(mm, 1.0f, s);
  for (int j = 0; j < 3; ++j) vol *= std::max(ext[j], emax * 1e-3);
  double c0 = std::max(std::cbrt(vol / (double)n_ * (double)k), emax * 1e-6);
  double c = c0;
  float b0[6];
  margin_bbox(mm, (float)c0, b0);
  const double g0 = cells_on_axis(mm[3], b0[0], (float)c0), g1 = cells_on_axis(mm[4], b0[1], (float)c0),
               g2 = cells_on_axis(mm[5], b0[2], (float)c0);
  if (g0 * g1 * g2 <= 4194304.0) {
    const size_t nc = (size_t)(g0 * g1 * g2);
    const size_t nw = ceil_div(nc, 4);
    Tmp flags, cnt;
    RVP_TRY(flags.alloc(nw * 4, s));
    RVP_TRY(cnt.alloc(sizeof(uint32_t), s));
    RVP_HIP(hipMemsetAsync(flags.p, 0, nw * 4, s));
    RVP_HIP(hipMemsetAsync(cnt.p, 0, sizeof(uint32_t), s));
    RVP_LAUNCH("occupancy_k", s, occupancy_k<<<ceil_div(n_, 256), 256, 0, s>>>(
                                     x_, y_, z_, n_, b0[0], b0[1], b0[2], (float)c0, (int)g0, (int)g1, (int)g2,
                                     flags.as<uint8_t>()));
    RVP_LAUNCH("count_marked_k", s, count_marked_k<<<(unsigned)std::min<size_t>(1024, ceil_div(nw, 256)), 256, 0, s>>>(
                                        flags.as<uint32_t>(), nw, cnt.as<uint32_t>()));
    uint32_t occupied0 = 0;
    RVP_HIP(hipMemcpyAsync(&occupied0, cnt.p, sizeof(occupied0), hipMemcpyDeviceToHost, s));
    RVP_HIP(hipStreamSynchronize(s));
    const double m = (double)n_ / (double)std::max<uint32_t>(occupied0, 1);
    const double R = c0 * std::sqrt((double)k / (3.14159265358979 * m));
    c = std::min(std::max(R / tune_knn_cell_div(), c0 / 64.0), c0 * 2.0);
  }
  return build_from(mm, (float)c, s);
}

GridView Grid::view() const {
  GridView v;
  v.xs = xs.as<float>();
  v.ys = ys.as<float>();
  v.zs = zs.as<float>();
  v.perm = perm.as<uint32_t>();
  v.rank = rank.as<uint32_t>();
  v.cell_start = cstart.as<uint32_t>();
  v.mnx = bb[0];
  v.mny = bb[1];
  v.mnz = bb[2];
  v.cell = cell;
  v.icell = 1.0f / cell;
  v.gx = dims[0];
  v.gy = dims[1];
  v.gz = dims[2];
  v.n = (uint32_t)n;
  return v;
}

// reference cell coordinates and hash keys (spatial_hashing.h:43-52) with the caller's cell size
__global__ void ref_cells_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                            size_t n, float mnx, float mny, float mnz, float c, int32_t *__restrict__ ixyz,
                            int64_t *__restrict__ keys) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int ix = (int)floorf(__fdiv_rn(__fsub_rn(x[i], mnx), c));
  int iy = (int)floorf(__fdiv_rn(__fsub_rn(y[i], mny), c));
  int iz = (int)floorf(__fdiv_rn(__fsub_rn(z[i], mnz), c));
  if (ixyz) {
    ixyz[3 * i + 0] = ix;
    ixyz[3 * i + 1] = iy;
    ixyz[3 * i + 2] = iz;
  }
  if (keys) keys[i] = (int64_t)ix + (int64_t)iy * 73856093LL + (int64_t)iz * 19349663LL;
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_hash_create(rvp_hash **h) {
  RVP_CHECK(h, RVP_E_ARG, "null handle pointer");
  RVP_TRY(pool_init());
  *h = new (std::nothrow) rvp_hash();
  RVP_CHECK(*h, RVP_E_OOM, "host allocation failed");
  (*h)->g.persistent = true;
  return RVP_OK;
}

int rvp_hash_destroy(rvp_hash *h) {
  if (!h) return RVP_OK;
  delete h;
  return RVP_OK;
}

int rvp_hash_build(rvp_hash *h, const float *x, const float *y, const float *z, size_t n, float cell,
                   rvp_stream_t s) {
  RVP_CHECK(h, RVP_E_ARG, "null handle");
  return h->g.build(x, y, z, n, cell, S(s));
}

int rvp_hash_bbox(const rvp_hash *h, float *bbox6) {
  RVP_CHECK(h && bbox6, RVP_E_ARG, "null argument");
  RVP_CHECK(h->g.built, RVP_E_STATE, "hash not built");
  std::copy(h->g.bb, h->g.bb + 6, bbox6);
  return RVP_OK;
}

int rvp_hash_info(const rvp_hash *h, size_t *n, size_t *occ, int *dims3, float *grid_cell) {
  RVP_CHECK(h, RVP_E_ARG, "null handle");
  RVP_CHECK(h->g.built, RVP_E_STATE, "hash not built");
  if (n) *n = h->g.n;
  if (occ) *occ = h->g.occupied_cells();
  if (dims3) std::copy(h->g.dims, h->g.dims + 3, dims3);
  if (grid_cell) *grid_cell = h->g.cell;
  return RVP_OK;
}

int rvp_hash_cell_coords(const rvp_hash *h, int32_t *ixyz, int64_t *keys, rvp_stream_t s) {
  RVP_CHECK(h, RVP_E_ARG, "null handle");
  RVP_CHECK(h->g.built, RVP_E_STATE, "hash not built");
  const Grid &g = h->g;
  if (g.n == 0) return RVP_OK;
  RVP_LAUNCH("ref_cells_k", S(s), ref_cells_k<<<ceil_div(g.n, 256), 256, 0, S(s)>>>(g.x, g.y, g.z, g.n, g.bb[0], g.bb[1], g.bb[2], g.req_cell,
                                                   ixyz, keys));
  RVP_LAUNCHED();
  RVP_HIP(hipStreamSynchronize(S(s)));
  return RVP_OK;
}

}  // extern "C"
