// voxel_ransac.hip -- V1 voxel downsampling, R1 RANSAC plane (GPU inlier counting + host replay of the
// adaptive stop), X1 plane inlier/outlier extraction.
#include <algorithm>
#include <climits>
#include <cmath>
#include <limits>
#include <vector>

#include "prims.hpp"

namespace rvp {

// ------------------------------------------------------------------------------------------------
// V1: voxel_grid_downsamp_sc (src/filters/voxel_grid_downsamp.cpp:183-213)
//   v = (int)floor(x * fl(1/leaf)) per axis; output in std::map order = lexicographic (vx, vy, vz);
//   per-voxel fp32 sums in input order; centroid = sum * fl(1 / count).
// A stable radix sort on the mixed-radix key ((vx-mx)*Ry + (vy-my))*Rz + (vz-mz) groups each voxel's
// points in input order, so one thread per voxel reproduces the sequential sums bit for bit.
// ------------------------------------------------------------------------------------------------
__global__ void voxel_minmax_init_k(int *mm) {
  if (threadIdx.x < 3) mm[threadIdx.x] = INT_MAX;
  else if (threadIdx.x < 6) mm[threadIdx.x] = INT_MIN;
}

__device__ __forceinline__ int vcoord(float v, float inv) { return (int)floorf(__fmul_rn(v, inv)); }

__global__ __launch_bounds__(256) void voxel_minmax_k(const float *__restrict__ x, const float *__restrict__ y,
                                                       const float *__restrict__ z, size_t n, float inv,
                                                       int *__restrict__ mm) {
  int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int a = vcoord(x[i], inv), b = vcoord(y[i], inv), c = vcoord(z[i], inv);
    v[0] = min(v[0], a); v[1] = min(v[1], b); v[2] = min(v[2], c);
    v[3] = max(v[3], a); v[4] = max(v[4], b); v[5] = max(v[5], c);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1)
    for (int j = 0; j < 3; ++j) {
      v[j] = min(v[j], __shfl_xor(v[j], d));
      v[j + 3] = max(v[j + 3], __shfl_xor(v[j + 3], d));
    }
  __shared__ int sm[4][6];
  if (lane() == 0)
    for (int j = 0; j < 6; ++j) sm[threadIdx.x >> 6][j] = v[j];
  __syncthreads();
  if (threadIdx.x < 3) {
    const int j = threadIdx.x;
    atomicMin(&mm[j], min(min(sm[0][j], sm[1][j]), min(sm[2][j], sm[3][j])));
    atomicMax(&mm[j + 3], max(max(sm[0][j + 3], sm[1][j + 3]), max(sm[2][j + 3], sm[3][j + 3])));
  }
}

template <class K>
__global__ void voxel_keys_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                             size_t n, float inv, int mx, int my, int mz, K ry, K rz, K *__restrict__ keys,
                             uint32_t *__restrict__ vals) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const K a = (K)(uint32_t)(vcoord(x[i], inv) - mx);
  const K b = (K)(uint32_t)(vcoord(y[i], inv) - my);
  const K c = (K)(uint32_t)(vcoord(z[i], inv) - mz);
  keys[i] = (a * ry + b) * rz + c;
  vals[i] = (uint32_t)i;
}

template <class K>
__global__ void run_flags_k(const K *__restrict__ keys, size_t n, uint32_t *__restrict__ f) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

template <class K>
__global__ void run_starts_k(const K *__restrict__ keys, size_t n, const uint32_t *__restrict__ vid,
                             uint32_t *__restrict__ start) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || keys[i] != keys[i - 1]) start[vid[i]] = (uint32_t)i;
}

__global__ void voxel_centroid_k(const float *__restrict__ x, const float *__restrict__ y,
                                 const float *__restrict__ z, const uint32_t *__restrict__ perm,
                                 const uint32_t *__restrict__ start, const uint32_t *__restrict__ nvox, uint32_t n,
                                 float *__restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t V = *nvox;
  if (v >= V) return;
  const uint32_t s = start[v], e = (v + 1 < V) ? start[v + 1] : n;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (uint32_t j = s; j < e; ++j) {
    const uint32_t p = perm[j];
    sx = __fadd_rn(sx, x[p]);
    sy = __fadd_rn(sy, y[p]);
    sz = __fadd_rn(sz, z[p]);
  }
  const float f = __fdiv_rn(1.0f, __uint2float_rn(e - s));
  out[3 * (size_t)v + 0] = __fmul_rn(sx, f);
  out[3 * (size_t)v + 1] = __fmul_rn(sy, f);
  out[3 * (size_t)v + 2] = __fmul_rn(sz, f);
}

template <class K>
static int voxel_sorted(const float *x, const float *y, const float *z, size_t n, float inv, const int mm[6],
                        uint64_t ry, uint64_t rz, int bits, float *out, size_t *n_out, hipStream_t s) {
  Tmp keys, vals, kt, vt, vid, start, nv;
  RVP_TRY(keys.alloc(n * sizeof(K), s));
  RVP_TRY(vals.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(kt.alloc(n * sizeof(K), s));
  RVP_TRY(vt.alloc(n * sizeof(uint32_t), s));
  const unsigned nb = ceil_div(n, 256);
  RVP_LAUNCH("voxel_keys_k", s, voxel_keys_k<K><<<nb, 256, 0, s>>>(x, y, z, n, inv, mm[0], mm[1], mm[2], (K)ry, (K)rz, keys.as<K>(),
                                     vals.as<uint32_t>()));
  RVP_LAUNCHED();
  K *kp = keys.as<K>(), *ktp = kt.as<K>();
  uint32_t *vp = vals.as<uint32_t>(), *vtp = vt.as<uint32_t>();
  if constexpr (sizeof(K) == 4) RVP_TRY(radix_sort_pairs_u32(&kp, &vp, &ktp, &vtp, n, bits, s));
  else RVP_TRY(radix_sort_pairs_u64(&kp, &vp, &ktp, &vtp, n, bits, s));
  RVP_TRY(vid.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(nv.alloc(sizeof(uint32_t), s));
  RVP_LAUNCH("run_flags_k", s, run_flags_k<K><<<nb, 256, 0, s>>>(kp, n, vid.as<uint32_t>()));
  RVP_TRY(exclusive_scan_u32(vid.as<uint32_t>(), vid.as<uint32_t>(), n, nv.as<uint32_t>(), s));
  RVP_TRY(start.alloc(n * sizeof(uint32_t), s));
  RVP_LAUNCH("run_starts_k", s, run_starts_k<K><<<nb, 256, 0, s>>>(kp, n, vid.as<uint32_t>(), start.as<uint32_t>()));
  RVP_LAUNCH("voxel_centroid_k", s, voxel_centroid_k<<<nb, 256, 0, s>>>(x, y, z, vp, start.as<uint32_t>(), nv.as<uint32_t>(), (uint32_t)n, out));
  uint32_t V = 0;
  RVP_HIP(hipMemcpyAsync(&V, nv.p, sizeof(V), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *n_out = V;
  return RVP_OK;
}

// ------------------------------------------------------------------------------------------------
// R1: hypotheses, plane coefficients (compute_plane_coefficients, ransac_plane.cpp:18-50), counting
// ------------------------------------------------------------------------------------------------

// glibc TYPE_3 random_r: r[i] = r[i-3] + r[i-31] (mod 2^32), output r >> 1, state seeded by the
// multiplier-16807 Lehmer generator (mod 2^31 - 1), first 310 outputs discarded.
struct GlibcRand {
  int32_t r[34];
  int pos = 0;
  explicit GlibcRand(unsigned seed) {
    if (seed == 0) seed = 1;
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
      int64_t w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 310; ++i) (void)next();
  }
  int32_t next() {
    const uint32_t v = (uint32_t)r[(pos + 3) % 34] + (uint32_t)r[(pos + 31) % 34];
    r[pos] = (int32_t)v;
    pos = (pos + 1) % 34;
    return (int32_t)(v >> 1);
  }
};

__global__ void plane_models_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                               const int32_t *__restrict__ idx3, int H, float collinear, float *__restrict__ models,
                               int32_t *__restrict__ valid) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= H) return;
  const int i1 = idx3[3 * h], i2 = idx3[3 * h + 1], i3 = idx3[3 * h + 2];
  if (i1 < 0) {  // duplicate draw: the iteration is consumed without a model (ransac_plane.cpp:72)
    valid[h] = 0;
    return;
  }
  const float x1 = x[i1], y1 = y[i1], z1 = z[i1];
  const float v1x = __fsub_rn(x[i2], x1), v1y = __fsub_rn(y[i2], y1), v1z = __fsub_rn(z[i2], z1);
  const float v2x = __fsub_rn(x[i3], x1), v2y = __fsub_rn(y[i3], y1), v2z = __fsub_rn(z[i3], z1);
  float a = __fsub_rn(__fmul_rn(v1y, v2z), __fmul_rn(v1z, v2y));
  float b = __fsub_rn(__fmul_rn(v1z, v2x), __fmul_rn(v1x, v2z));
  float c = __fsub_rn(__fmul_rn(v1x, v2y), __fmul_rn(v1y, v2x));
  const float norm = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(c, c)));
  if (norm < collinear) {
    valid[h] = 0;
    return;
  }
  a = __fdiv_rn(a, norm);
  b = __fdiv_rn(b, norm);
  c = __fdiv_rn(c, norm);
  const float d = -__fadd_rn(__fadd_rn(__fmul_rn(a, x1), __fmul_rn(b, y1)), __fmul_rn(c, z1));
  models[4 * h + 0] = a;
  models[4 * h + 1] = b;
  models[4 * h + 2] = c;
  models[4 * h + 3] = d;
  valid[h] = 1;
}

constexpr int RC_PTS = 4;       // points per thread
constexpr int RC_HYP = 32;      // hypotheses per block
constexpr int RC_TILE = 256 * RC_PTS;

// inliers: fabsf(((a*x + b*y) + c*z) + d) <= thr  (ransac_plane.cpp:81-89)
__global__ __launch_bounds__(256) void plane_count_k(const float *__restrict__ x, const float *__restrict__ y,
                                                      const float *__restrict__ z, uint32_t n,
                                                      const float *__restrict__ models,
                                                      const int32_t *__restrict__ valid, int H, float thr,
                                                      int32_t *__restrict__ counts) {
  __shared__ int sc[RC_HYP];
  if (threadIdx.x < RC_HYP) sc[threadIdx.x] = 0;
  __syncthreads();
  float px[RC_PTS], py[RC_PTS], pz[RC_PTS];
  bool ok[RC_PTS];
#pragma unroll
  for (int j = 0; j < RC_PTS; ++j) {
    const uint32_t i = blockIdx.x * RC_TILE + j * 256 + threadIdx.x;
    ok[j] = i < n;
    px[j] = ok[j] ? x[i] : 0.f;
    py[j] = ok[j] ? y[i] : 0.f;
    pz[j] = ok[j] ? z[i] : 0.f;
  }
  const int h0 = blockIdx.y * RC_HYP;
  const int hn = min(RC_HYP, H - h0);
  for (int hh = 0; hh < hn; ++hh) {
    const int h = h0 + hh;
    if (!valid[h]) continue;
    const float a = models[4 * h], b = models[4 * h + 1], c = models[4 * h + 2], d = models[4 * h + 3];
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < RC_PTS; ++j) {
      const float dist = fabsf(__fadd_rn(
          __fadd_rn(__fadd_rn(__fmul_rn(a, px[j]), __fmul_rn(b, py[j])), __fmul_rn(c, pz[j])), d));
      cnt += popc(ballot(ok[j] && dist <= thr));
    }
    if (lane() == 0 && cnt) atomicAdd(&sc[hh], cnt);
  }
  __syncthreads();
  if (threadIdx.x < hn && sc[threadIdx.x]) atomicAdd(&counts[h0 + threadIdx.x], sc[threadIdx.x]);
}

static int plane_counts_dev(const float *x, const float *y, const float *z, size_t n, const float *models,
                            const int32_t *valid, int H, float thr, int32_t *counts, hipStream_t s) {
  RVP_HIP(hipMemsetAsync(counts, 0, (size_t)H * sizeof(int32_t), s));
  if (n == 0 || H == 0) return RVP_OK;
  dim3 grid(ceil_div(n, RC_TILE), ceil_div((size_t)H, RC_HYP));
  RVP_LAUNCH("plane_count_k", s, plane_count_k<<<grid, 256, 0, s>>>(x, y, z, (uint32_t)n, models, valid, H, thr, counts));
  return RVP_OK;
}

// ------------------------------------------------------------------------------------------------
// X1: dist = ((a*x + b*y) + c*z) + d (no fabs); inlier iff dist >= -t && dist <= t (ransac_plane.cpp:374-382)
// ------------------------------------------------------------------------------------------------
__global__ void plane_flags_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                              size_t n, float a, float b, float c, float d, float t, uint8_t *__restrict__ f) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float dist = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(a, x[i]), __fmul_rn(b, y[i])), __fmul_rn(c, z[i])), d);
  f[i] = (dist >= -t && dist <= t) ? 1 : 0;
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_voxel(const float *x, const float *y, const float *z, size_t n, float leaf, float *out_aos, size_t *n_out,
              rvp_stream_t st) {
  RVP_CHECK(n_out, RVP_E_ARG, "null n_out");
  *n_out = 0;
  if (n == 0) return RVP_OK;
  RVP_CHECK(x && y && z && out_aos, RVP_E_ARG, "null pointer");
  RVP_CHECK(n < 0xfffffff0ull, RVP_E_UNSUPPORTED, "more than 2^32-16 points");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  const float inv = 1.0f / leaf;  // voxel_grid_downsamp.cpp:189 (IEEE division on the host)
  Tmp mmd;
  RVP_TRY(mmd.alloc(6 * sizeof(int), s));
  RVP_LAUNCH("voxel_minmax_init_k", s, voxel_minmax_init_k<<<1, 64, 0, s>>>(mmd.as<int>()));
  const int nb = (int)std::min<size_t>(256, ceil_div(n, 256 * 8));
  RVP_LAUNCH("voxel_minmax_k", s, voxel_minmax_k<<<nb, 256, 0, s>>>(x, y, z, n, inv, mmd.as<int>()));
  int mm[6];
  RVP_HIP(hipMemcpyAsync(mm, mmd.p, sizeof(mm), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  const uint64_t rx = (uint64_t)((int64_t)mm[3] - mm[0] + 1), ry = (uint64_t)((int64_t)mm[4] - mm[1] + 1),
                 rz = (uint64_t)((int64_t)mm[5] - mm[2] + 1);
  const long double cells = (long double)rx * (long double)ry * (long double)rz;
  RVP_CHECK(cells < 1.8e19L, RVP_E_UNSUPPORTED, "voxel key range exceeds 64 bits");
  const uint64_t maxkey = rx * ry * rz - 1;
  const int bits = bits_for(maxkey);
  if (maxkey <= 0xffffffffull)
    return voxel_sorted<uint32_t>(x, y, z, n, inv, mm, ry, rz, bits, out_aos, n_out, s);
  return voxel_sorted<uint64_t>(x, y, z, n, inv, mm, ry, rz, bits, out_aos, n_out, s);
}

int rvp_plane_counts(const float *x, const float *y, const float *z, size_t n, const float *models, int H,
                     float dist_thresh, int32_t *counts, rvp_stream_t st) {
  RVP_CHECK(H >= 0 && (H == 0 || (models && counts)), RVP_E_ARG, "bad models/counts");
  RVP_CHECK(n < 0xfffffff0ull, RVP_E_UNSUPPORTED, "size");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  Tmp valid;
  RVP_TRY(valid.alloc((size_t)std::max(H, 1) * sizeof(int32_t), s));
  std::vector<int32_t> ones((size_t)std::max(H, 1), 1);
  RVP_HIP(hipMemcpyAsync(valid.p, ones.data(), ones.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  RVP_TRY(plane_counts_dev(x, y, z, n, models, valid.as<int32_t>(), H, dist_thresh, counts, s));
  RVP_HIP(hipStreamSynchronize(s));
  return RVP_OK;
}

int rvp_ransac_plane(const float *x, const float *y, const float *z, size_t n, float dist_thresh, int max_iters,
                     float collinear_thresh, float probability, float *model, int *inliers, int *iterations,
                     rvp_stream_t st) {
  RVP_CHECK(model && inliers, RVP_E_ARG, "null output");
  *inliers = 0;
  if (iterations) *iterations = 0;
  for (int k = 0; k < 4; ++k) model[k] = 0.0f;
  if (n < 3) return RVP_OK;  // ransac_plane.cpp:59
  RVP_CHECK(n < 0xfffffff0ull, RVP_E_UNSUPPORTED, "size");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);

  GlibcRand rng(0);  // std::srand(0) (ransac_plane.cpp:60)
  int best = 0;
  float best_model[4] = {0, 0, 0, 0};
  int k_iters = max_iters;
  const double log_p = std::log(1.0 - std::clamp((double)probability, 0.5, 0.9999));

  Tmp idx_d, models_d, valid_d, counts_d;
  int iter = 0;
  int chunk = 128;
  bool done = max_iters <= 0;
  while (!done) {
    const int H = std::min(chunk, max_iters - iter);
    if (H <= 0) break;
    // three draws per iteration whether or not it is skipped (ransac_plane.cpp:69-72)
    std::vector<int32_t> idx3((size_t)H * 3);
    for (int h = 0; h < H; ++h) {
      const int i1 = (int)((size_t)rng.next() % n);
      const int i2 = (int)((size_t)rng.next() % n);
      const int i3 = (int)((size_t)rng.next() % n);
      const bool dup = i1 == i2 || i1 == i3 || i2 == i3;
      idx3[3 * h] = dup ? -1 : i1;
      idx3[3 * h + 1] = i2;
      idx3[3 * h + 2] = i3;
    }
    RVP_TRY(idx_d.alloc(idx3.size() * sizeof(int32_t), s));
    RVP_TRY(models_d.alloc((size_t)H * 4 * sizeof(float), s));
    RVP_TRY(valid_d.alloc((size_t)H * sizeof(int32_t), s));
    RVP_TRY(counts_d.alloc((size_t)H * sizeof(int32_t), s));
    RVP_HIP(hipMemcpyAsync(idx_d.p, idx3.data(), idx3.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    RVP_LAUNCH("plane_models_k", s, plane_models_k<<<ceil_div(H, 64), 64, 0, s>>>(x, y, z, idx_d.as<int32_t>(), H, collinear_thresh,
                                                  models_d.as<float>(), valid_d.as<int32_t>()));
    RVP_LAUNCHED();
    RVP_TRY(plane_counts_dev(x, y, z, n, models_d.as<float>(), valid_d.as<int32_t>(), H, dist_thresh,
                             counts_d.as<int32_t>(), s));
    std::vector<float> models((size_t)H * 4);
    std::vector<int32_t> valid(H), counts(H);
    RVP_HIP(hipMemcpyAsync(models.data(), models_d.p, models.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    RVP_HIP(hipMemcpyAsync(valid.data(), valid_d.p, valid.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RVP_HIP(hipMemcpyAsync(counts.data(), counts_d.p, counts.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RVP_HIP(hipStreamSynchronize(s));
    // replay of the sequential loop (ransac_plane.cpp:68-107)
    for (int h = 0; h < H; ++h, ++iter) {
      if (!(iter < k_iters && iter < max_iters)) {
        done = true;
        break;
      }
      if (!valid[h]) continue;
      const int cur = counts[h];
      if (cur > best) {
        best = cur;
        for (int k = 0; k < 4; ++k) best_model[k] = models[4 * h + k];
        const double w = (double)best / (double)n;
        double p_no = 1.0 - std::pow(w, 3.0);
        p_no = std::max(std::numeric_limits<double>::epsilon(), p_no);
        p_no = std::min(1.0 - std::numeric_limits<double>::epsilon(), p_no);
        const double log_no = std::log(p_no);
        if (std::abs(log_no) > 1e-7) {
          const int dyn = (int)std::ceil(log_p / log_no);
          if (dyn > 0 && dyn < k_iters) k_iters = dyn;
        }
      }
    }
    if (iter >= k_iters || iter >= max_iters) done = true;
    chunk = std::min(chunk * 2, 1024);
  }
  for (int k = 0; k < 4; ++k) model[k] = best_model[k];
  *inliers = best;
  if (iterations) *iterations = std::min(iter, std::min(k_iters, max_iters));
  return RVP_OK;
}

int rvp_plane_extract(const float *x, const float *y, const float *z, size_t n, const float *m, float t, float *in_aos,
                      size_t *n_in, float *out_aos, size_t *n_out, rvp_stream_t st) {
  RVP_CHECK(m && n_in && n_out, RVP_E_ARG, "null argument");
  *n_in = *n_out = 0;
  if (n == 0) return RVP_OK;
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  Tmp f, c;
  RVP_TRY(f.alloc(n, s));
  RVP_TRY(c.alloc(2 * sizeof(uint32_t), s));
  RVP_LAUNCH("plane_flags_k", s, plane_flags_k<<<ceil_div(n, 256), 256, 0, s>>>(x, y, z, n, m[0], m[1], m[2], m[3], t, f.as<uint8_t>()));
  RVP_TRY(compact_xyz_aos(x, y, z, f.as<uint8_t>(), 1, n, in_aos, c.as<uint32_t>(), s));
  RVP_TRY(compact_xyz_aos(x, y, z, f.as<uint8_t>(), 0, n, out_aos, c.as<uint32_t>() + 1, s));
  uint32_t h[2];
  RVP_HIP(hipMemcpyAsync(h, c.p, sizeof(h), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *n_in = h[0];
  *n_out = h[1];
  return RVP_OK;
}

}  // extern "C"
