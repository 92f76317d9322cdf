// voxel_ransac.hip -- V1 voxel downsampling, R1 RANSAC plane (GPU inlier counting + host replay of the
// adaptive stop), X1 plane inlier/outlier extraction.
#include <algorithm>
#include <climits>
#include <cmath>
#include <limits>
#include <mutex>
#include <vector>

#include "batch.hpp"
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
      v[j] = min(v[j], shfl_xor(v[j], d));
      v[j + 3] = max(v[j + 3], shfl_xor(v[j + 3], d));
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

__global__ void voxel_centroid_k(const float *__restrict__ x, const float *__restrict__ y,
                                 const float *__restrict__ z, const uint32_t *__restrict__ perm,
                                 const uint32_t *__restrict__ start, const uint32_t *__restrict__ nvox, uint32_t n_host,
                                 const uint32_t *__restrict__ n_dev, float *__restrict__ out) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t V = *nvox;
  const uint32_t n = n_dev ? min(*n_dev, n_host) : n_host;
  if (v >= V) return;
  const uint32_t s = start[v], e = (v + 1 < V) ? start[v + 1] : n;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  uint32_t j = s;
  // the loads run ahead of the sums (8 indices, then their coordinates); the adds stay in input order
  for (; j + 8 <= e; j += 8) {
    uint32_t p[8];
    float vx[8], vy[8], vz[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] = perm[j + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      vx[u] = x[p[u]];
      vy[u] = y[p[u]];
      vz[u] = z[p[u]];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      sx = __fadd_rn(sx, vx[u]);
      sy = __fadd_rn(sy, vy[u]);
      sz = __fadd_rn(sz, vz[u]);
    }
  }
  for (; j < e; ++j) {
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
  const bool large = n >= (1u << 19);
  if constexpr (sizeof(K) == 4) RVP_TRY(radix_sort_pairs_u32(&kp, &vp, &ktp, &vtp, n, bits, s, nullptr, nullptr, large));
  else RVP_TRY(radix_sort_pairs_u64(&kp, &vp, &ktp, &vtp, n, bits, s, nullptr, nullptr, large));
  RVP_TRY(vid.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(nv.alloc(sizeof(uint32_t), s));
  RVP_TRY(start.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(run_heads(kp, n, nullptr, vid.as<uint32_t>(), start.as<uint32_t>(), nv.as<uint32_t>(), s));
  RVP_LAUNCH("voxel_centroid_k", s, voxel_centroid_k<<<nb, 256, 0, s>>>(x, y, z, vp, start.as<uint32_t>(),
                                                                        nv.as<uint32_t>(), (uint32_t)n, nullptr, out));
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
  // the block's models in LDS first: the loop below then reads no global memory
  __shared__ float4 sm[RC_HYP];
  __shared__ int sv[RC_HYP];
  if (threadIdx.x < hn) {
    const int h = h0 + threadIdx.x;
    sm[threadIdx.x] = make_float4(models[4 * h], models[4 * h + 1], models[4 * h + 2], models[4 * h + 3]);
    sv[threadIdx.x] = valid[h];
  }
  __syncthreads();
  for (int hh = 0; hh < hn; ++hh) {
    if (!sv[hh]) continue;
    const float4 m = sm[hh];
    const float a = m.x, b = m.y, c = m.z, d = m.w;
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

// Replay of the reference's sequential loop (ransac_plane.cpp:68-107) over counted hypotheses, fed in
// order; the state carries across feeds.  Host fp64 log / pow as the reference's.
struct RansacReplay {
  int best = 0;
  float model[4] = {0, 0, 0, 0};
  int k_iters, max_iters, iter = 0;
  size_t n;
  double log_p;
  bool done;
  RansacReplay(size_t n_, int max_iters_, float probability)
      : k_iters(max_iters_), max_iters(max_iters_), n(n_),
        log_p(std::log(1.0 - std::clamp((double)probability, 0.5, 0.9999))), done(max_iters_ <= 0) {}
  void feed(const float *models, const int32_t *valid, const int32_t *counts, int H) {
    for (int h = 0; h < H && !done; ++h, ++iter) {
      if (!(iter < k_iters && iter < max_iters)) {
        done = true;
        break;
      }
      if (!valid[h]) continue;
      const int cur = counts[h];
      if (cur > best) {
        best = cur;
        for (int k = 0; k < 4; ++k) model[k] = models[4 * h + k];
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
  }
  int iterations() const { return iter; }  // iterations the reference's loop runs
};

static int plane_counts_dev(const float *x, const float *y, const float *z, size_t n, const float *models,
                            const int32_t *valid, int H, float thr, int32_t *counts, hipStream_t s) {
  RVP_HIP(hipMemsetAsync(counts, 0, (size_t)H * sizeof(int32_t), s));
  if (n == 0 || H == 0) return RVP_OK;
  dim3 grid(ceil_div(n, RC_TILE), ceil_div((size_t)H, RC_HYP));
  RVP_LAUNCH("plane_count_k", s, plane_count_k<<<grid, 256, 0, s>>>(x, y, z, (uint32_t)n, models, valid, H, thr, counts));
  return RVP_OK;
}

// ------------------------------------------------------------------------------------------------
// batched V1 and R1: clouds [off[c], off[c+1]) of one SoA batch, counts on the device, no host sync
// except the final read of the RANSAC counts.
// ------------------------------------------------------------------------------------------------
// voxel coordinate bounds per cloud (blockIdx.y = cloud), one partial per block in pmm[c][block][6];
// inv = fl(1 / leaf) as voxel_grid_downsamp.cpp:189
__global__ __launch_bounds__(256) void seg_vox_minmax_k(const float *__restrict__ x, const float *__restrict__ y,
                                                         const float *__restrict__ z, const uint32_t *__restrict__ off,
                                                         const float *__restrict__ leaf, int *__restrict__ pmm) {
  const int c = blockIdx.y;
  const float inv = __fdiv_rn(1.0f, leaf[c]);
  int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  const uint32_t e = off[c + 1];
  for (uint32_t i = off[c] + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    const int a = vcoord(x[i], inv), b = vcoord(y[i], inv), cc = vcoord(z[i], inv);
    v[0] = min(v[0], a); v[1] = min(v[1], b); v[2] = min(v[2], cc);
    v[3] = max(v[3], a); v[4] = max(v[4], b); v[5] = max(v[5], cc);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1)
    for (int j = 0; j < 3; ++j) {
      v[j] = min(v[j], shfl_xor(v[j], d));
      v[j + 3] = max(v[j + 3], shfl_xor(v[j + 3], d));
    }
  __shared__ int sm[4][6];
  if (lane() == 0)
    for (int j = 0; j < 6; ++j) sm[threadIdx.x >> 6][j] = v[j];
  __syncthreads();
  if (threadIdx.x < 3) {
    const int j = threadIdx.x;
    int *p = pmm + ((size_t)c * gridDim.x + blockIdx.x) * 6;
    p[j] = min(min(sm[0][j], sm[1][j]), min(sm[2][j], sm[3][j]));
    p[j + 3] = max(max(sm[0][j + 3], sm[1][j + 3]), max(sm[2][j + 3], sm[3][j + 3]));
  }
}

// the per-block partials of cloud c (blockIdx.x) -> vmm[c][6]
__device__ __forceinline__ void seg_vox_reduce_cloud(const int *__restrict__ pmm, int per, int c, int *__restrict__ vmm) {
  const int l = lane();
  int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  for (int b = l; b < per; b += WAVE) {
    const int *p = pmm + ((size_t)c * per + b) * 6;
    for (int j = 0; j < 3; ++j) {
      v[j] = min(v[j], p[j]);
      v[j + 3] = max(v[j + 3], p[j + 3]);
    }
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1)
    for (int j = 0; j < 3; ++j) {
      v[j] = min(v[j], shfl_xor(v[j], d));
      v[j + 3] = max(v[j + 3], shfl_xor(v[j + 3], d));
    }
  if (l == 0)
    for (int j = 0; j < 6; ++j) vmm[6 * c + j] = v[j];
}

__global__ __launch_bounds__(64) void seg_vox_reduce_k(const int *__restrict__ pmm, int per, int *__restrict__ vmm) {
  seg_vox_reduce_cloud(pmm, per, blockIdx.x, vmm);
}

// key space of each cloud: rx * ry * rz voxels (1 for an empty cloud or one whose range overflows 2^48,
// which is then flagged in status[c]); base[c] = exclusive prefix (one wave); *bits = key width
__device__ __forceinline__ void seg_vox_base_wave(int ncl, const uint32_t *__restrict__ off,
                                                     const int *__restrict__ vmm, uint64_t *__restrict__ ryz,
                                                     uint64_t *__restrict__ base, int *__restrict__ bits,
                                                     uint32_t *__restrict__ status) {
  const int l = lane();
  uint64_t carry = 0;
  for (int b0 = 0; b0 < ncl; b0 += 64) {
    const int c = b0 + l;
    uint64_t r = 0;
    if (c < ncl) {
      uint32_t st = 0;
      uint64_t ry = 1, rz = 1;
      r = 1;
      if (off[c + 1] > off[c]) {
        const int *m = vmm + 6 * c;
        const uint64_t rx = (uint64_t)((int64_t)m[3] - m[0] + 1);
        ry = (uint64_t)((int64_t)m[4] - m[1] + 1);
        rz = (uint64_t)((int64_t)m[5] - m[2] + 1);
        const double cells = (double)rx * (double)ry * (double)rz;
        if (cells < 0x1p48) r = rx * ry * rz;
        else st = 1;
      }
      ryz[2 * c] = ry;
      ryz[2 * c + 1] = rz;
      status[c] = st;
    }
    uint64_t inc = r;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint64_t t = __shfl_up(inc, d);
      if (l >= d) inc += t;
    }
    if (c < ncl) base[c] = carry + inc - r;
    carry += __shfl(inc, 63);
  }
  if (l == 0) {
    base[ncl] = carry;
    int b = 0;
    const uint64_t mx = carry ? carry - 1 : 0;
    while (b < 64 && (mx >> b) != 0) ++b;
    *bits = b;
  }
}

__global__ __launch_bounds__(64) void seg_vox_base_k(int ncl, const uint32_t *__restrict__ off,
                                                     const int *__restrict__ vmm, uint64_t *__restrict__ ryz,
                                                     uint64_t *__restrict__ base, int *__restrict__ bits,
                                                     uint32_t *__restrict__ status) {
  seg_vox_base_wave(ncl, off, vmm, ryz, base, bits, status);
}

// Both steps in one block, for batches of up to VOX_BOUNDS_MAX clouds: each wave reduces the partials of every
// 16th cloud, then wave 0 lays out the key space.
constexpr int VOX_BOUNDS_MAX = 4096;
__global__ __launch_bounds__(1024) void seg_vox_bounds_k(const int *__restrict__ pmm, int per, int ncl,
                                                         const uint32_t *__restrict__ off, int *__restrict__ vmm,
                                                         uint64_t *__restrict__ ryz, uint64_t *__restrict__ base,
                                                         int *__restrict__ bits, uint32_t *__restrict__ status) {
  const int w = wave_in_block(), nw = (int)(blockDim.x / WAVE);
  for (int c = w; c < ncl; c += nw) seg_vox_reduce_cloud(pmm, per, c, vmm);
  __syncthreads();
  if (w == 0) seg_vox_base_wave(ncl, off, vmm, ryz, base, bits, status);
}

__global__ void seg_vox_keys_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                               const uint32_t *__restrict__ off, int ncl, const float *__restrict__ leaf,
                               const int *__restrict__ vmm, const uint64_t *__restrict__ ryz,
                               const uint64_t *__restrict__ base, const uint32_t *__restrict__ status,
                               uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= off[ncl]) return;
  const int c = cloud_of(off, ncl, i);
  uint64_t k = base[c];
  if (!status[c]) {
    const float inv = __fdiv_rn(1.0f, leaf[c]);
    const int *m = vmm + 6 * c;
    const uint64_t a = (uint64_t)(uint32_t)(vcoord(x[i], inv) - m[0]);
    const uint64_t b = (uint64_t)(uint32_t)(vcoord(y[i], inv) - m[1]);
    const uint64_t cc = (uint64_t)(uint32_t)(vcoord(z[i], inv) - m[2]);
    k += (a * ryz[2 * c] + b) * ryz[2 * c + 1] + cc;
  }
  keys[i] = k;
  vals[i] = i;
}

// first output voxel of each cloud: the run index at its first sorted point (clouds are key-major)
__global__ void seg_vox_offsets_k(const uint32_t *__restrict__ off, int ncl, const uint32_t *__restrict__ vid,
                                  const uint32_t *__restrict__ nvox, uint32_t *__restrict__ out_off) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncl) return;
  const uint32_t o = off[c];
  out_off[c] = o < off[ncl] ? vid[o] : *nvox;
}

// Tiles of RC_TILE points that never straddle clouds: tile_off[c] = first tile of cloud c; also copies the
// offsets into the read-back buffer and marks every cloud active for the first round of counts.  One wave.
__device__ __forceinline__ void seg_tiles_wave(const uint32_t *__restrict__ off, int ncl, uint32_t *__restrict__ tile_off,
                                               uint32_t *__restrict__ off_copy, uint8_t *__restrict__ active) {
  const int l = lane();
  uint32_t carry = 0;
  for (int b0 = 0; b0 < ncl; b0 += 64) {
    const int c = b0 + l;
    const uint32_t t = c < ncl ? (off[c + 1] - off[c] + RC_TILE - 1) / RC_TILE : 0u;
    const uint32_t inc = wave_incl_scan(t);
    if (c < ncl) {
      tile_off[c] = carry + inc - t;
      off_copy[c] = off[c];
      active[c] = 1;
    }
    carry += __shfl(inc, 63);
  }
  if (l == 0) {
    tile_off[ncl] = carry;
    off_copy[ncl] = off[ncl];
  }
}

// hypotheses of every cloud from one rand() stream (each cloud's call starts with srand(0)):
// draws r[3h..3h+2] % n_c (ransac_plane.cpp:69-71), model as compute_plane_coefficients
__global__ void seg_plane_models_k(const float *__restrict__ x, const float *__restrict__ y,
                                   const float *__restrict__ z, const uint32_t *__restrict__ off, int ncl,
                                   const int32_t *__restrict__ draws, int H, float collinear,
                                   float *__restrict__ models, int32_t *__restrict__ valid,
                                   int32_t *__restrict__ counts, uint32_t *__restrict__ tile_off,
                                   uint32_t *__restrict__ off_copy, uint8_t *__restrict__ active) {
  if (blockIdx.x == 0 && threadIdx.x < WAVE) seg_tiles_wave(off, ncl, tile_off, off_copy, active);
  const size_t gh = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gh >= (size_t)ncl * H) return;
  const int c = (int)(gh / H), h = (int)(gh % H);
  const uint32_t n = off[c + 1] - off[c];
  valid[gh] = 0;
  counts[gh] = 0;  // the count kernel adds into it
  if (n < 3) return;
  const uint32_t i1 = (uint32_t)draws[3 * h] % n, i2 = (uint32_t)draws[3 * h + 1] % n,
                 i3 = (uint32_t)draws[3 * h + 2] % n;
  if (i1 == i2 || i1 == i3 || i2 == i3) return;
  const uint32_t o = off[c];
  const float x1 = x[o + i1], y1 = y[o + i1], z1 = z[o + i1];
  const float v1x = __fsub_rn(x[o + i2], x1), v1y = __fsub_rn(y[o + i2], y1), v1z = __fsub_rn(z[o + i2], z1);
  const float v2x = __fsub_rn(x[o + i3], x1), v2y = __fsub_rn(y[o + i3], y1), v2z = __fsub_rn(z[o + i3], z1);
  float a = __fsub_rn(__fmul_rn(v1y, v2z), __fmul_rn(v1z, v2y));
  float b = __fsub_rn(__fmul_rn(v1z, v2x), __fmul_rn(v1x, v2z));
  float cc = __fsub_rn(__fmul_rn(v1x, v2y), __fmul_rn(v1y, v2x));
  const float norm = sqrtf(__fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(cc, cc)));
  if (norm < collinear) return;
  a = __fdiv_rn(a, norm);
  b = __fdiv_rn(b, norm);
  cc = __fdiv_rn(cc, norm);
  const float d = -__fadd_rn(__fadd_rn(__fmul_rn(a, x1), __fmul_rn(b, y1)), __fmul_rn(cc, z1));
  models[4 * gh + 0] = a;
  models[4 * gh + 1] = b;
  models[4 * gh + 2] = cc;
  models[4 * gh + 3] = d;
  valid[gh] = 1;
}

// hypotheses [h_begin, h_end) of every cloud still active (active[c] != 0)
__global__ __launch_bounds__(256) void seg_plane_count_k(const float *__restrict__ x, const float *__restrict__ y,
                                                          const float *__restrict__ z, const uint32_t *__restrict__ off,
                                                          int ncl, const uint32_t *__restrict__ tile_off,
                                                          const float *__restrict__ models,
                                                          const int32_t *__restrict__ valid, int H, int h_begin,
                                                          int h_end, const uint8_t *__restrict__ active, float thr,
                                                          int32_t *__restrict__ counts) {
  const uint32_t tile = blockIdx.x;
  if (tile >= tile_off[ncl]) return;
  const int c = cloud_of(tile_off, ncl, tile);
  if (!active[c]) return;
  __shared__ int sc[RC_HYP];
  if (threadIdx.x < RC_HYP) sc[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t first = off[c] + (tile - tile_off[c]) * RC_TILE, end = off[c + 1];
  float px[RC_PTS], py[RC_PTS], pz[RC_PTS];
  bool ok[RC_PTS];
#pragma unroll
  for (int j = 0; j < RC_PTS; ++j) {
    const uint32_t i = first + j * 256 + threadIdx.x;
    ok[j] = i < end;
    px[j] = ok[j] ? x[i] : 0.f;
    py[j] = ok[j] ? y[i] : 0.f;
    pz[j] = ok[j] ? z[i] : 0.f;
  }
  const int h0 = h_begin + (int)blockIdx.y * RC_HYP;
  const int hn = min(RC_HYP, h_end - h0);
  const size_t mb = (size_t)c * H;
  // the block's models in LDS first: the loop below then reads no global memory
  __shared__ float4 sm[RC_HYP];
  __shared__ int sv[RC_HYP];
  if (threadIdx.x < hn) {
    const size_t h = mb + h0 + threadIdx.x;
    sm[threadIdx.x] = make_float4(models[4 * h], models[4 * h + 1], models[4 * h + 2], models[4 * h + 3]);
    sv[threadIdx.x] = valid[h];
  }
  __syncthreads();
  for (int hh = 0; hh < hn; ++hh) {
    if (!sv[hh]) continue;
    const float4 m = sm[hh];
    const float a = m.x, b = m.y, cc = m.z, d = m.w;
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < RC_PTS; ++j) {
      const float dist = fabsf(__fadd_rn(
          __fadd_rn(__fadd_rn(__fmul_rn(a, px[j]), __fmul_rn(b, py[j])), __fmul_rn(cc, pz[j])), d));
      cnt += popc(ballot(ok[j] && dist <= thr));
    }
    if (lane() == 0 && cnt) atomicAdd(&sc[hh], cnt);
  }
  __syncthreads();
  if (threadIdx.x < hn && sc[threadIdx.x]) atomicAdd(&counts[mb + h0 + threadIdx.x], sc[threadIdx.x]);
}

// glibc rand() after srand(0), cached per device and grown on demand
static int32_t *g_draws_dev[64] = {nullptr};
static size_t g_draws_len[64] = {0};

static std::mutex g_draws_mu;

// The table is copied on the caller's stream and waited for before it is published, so every launch that
// receives the pointer (on any stream) reads it complete; a plain hipMemcpy from pageable memory may return
// before the copy has landed, and the null stream it uses does not order against non-blocking streams.
static int rand_draws(size_t count, hipStream_t s, const int32_t **out) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  RVP_CHECK(dev >= 0 && dev < 64, RVP_E_UNSUPPORTED, "device index");
  std::lock_guard<std::mutex> lk(g_draws_mu);
  if (g_draws_len[dev] < count) {
    const size_t len = std::max<size_t>(count, 3072);
    std::vector<int32_t> h(len);
    GlibcRand rng(0);
    for (auto &v : h) v = rng.next();
    int32_t *d = nullptr;
    RVP_HIP(hipMalloc(&d, len * sizeof(int32_t)));
    RVP_HIP(hipMemcpyAsync(d, h.data(), len * sizeof(int32_t), hipMemcpyHostToDevice, s));
    RVP_HIP(hipStreamSynchronize(s));
    // a shorter previous table is kept: launches on other streams may still read it
    g_draws_dev[dev] = d;
    g_draws_len[dev] = len;
  }
  *out = g_draws_dev[dev];
  return RVP_OK;
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
  RansacReplay rp(n, max_iters, probability);
  // per chunk of H: the draws go up in one copy, models [H][4] + valid [H] + counts [H] come back in one,
  // both through pinned staging (draws in words [0, 3H), results after them)
  Tmp idx_d, pack;
  PinnedStage stage;
  int32_t *host = nullptr;
  RVP_TRY(stage.get(s, (size_t)1024 * 9 * sizeof(int32_t), reinterpret_cast<void **>(&host)));
  int chunk = 128, drawn = 0;
  while (!rp.done) {
    const int H = std::min(chunk, max_iters - drawn);
    if (H <= 0) break;
    // three draws per iteration whether or not it is skipped (ransac_plane.cpp:69-72)
    int32_t *idx3 = host;
    for (int h = 0; h < H; ++h) {
      const int i1 = (int)((size_t)rng.next() % n);
      const int i2 = (int)((size_t)rng.next() % n);
      const int i3 = (int)((size_t)rng.next() % n);
      const bool dup = i1 == i2 || i1 == i3 || i2 == i3;
      idx3[3 * h] = dup ? -1 : i1;
      idx3[3 * h + 1] = i2;
      idx3[3 * h + 2] = i3;
    }
    drawn += H;
    RVP_TRY(idx_d.alloc((size_t)H * 3 * sizeof(int32_t), s));
    RVP_TRY(pack.alloc((size_t)H * 6 * sizeof(int32_t), s));
    float *models_d = pack.as<float>();
    int32_t *valid_d = reinterpret_cast<int32_t *>(models_d + (size_t)H * 4), *counts_d = valid_d + H;
    RVP_HIP(hipMemcpyAsync(idx_d.p, idx3, (size_t)H * 3 * sizeof(int32_t), hipMemcpyHostToDevice, s));
    RVP_LAUNCH("plane_models_k", s, plane_models_k<<<ceil_div(H, 64), 64, 0, s>>>(x, y, z, idx_d.as<int32_t>(), H,
                                                                                  collinear_thresh, models_d, valid_d));
    RVP_TRY(plane_counts_dev(x, y, z, n, models_d, valid_d, H, dist_thresh, counts_d, s));
    int32_t *res = host + (size_t)H * 3;
    RVP_HIP(hipMemcpyAsync(res, pack.p, (size_t)H * 6 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RVP_HIP(hipStreamSynchronize(s));
    const float *hm = reinterpret_cast<const float *>(res);
    rp.feed(hm, res + (size_t)H * 4, res + (size_t)H * 5, H);
    chunk = std::min(chunk * 2, 1024);
  }
  for (int k = 0; k < 4; ++k) model[k] = rp.model[k];
  *inliers = rp.best;
  if (iterations) *iterations = rp.iterations();
  return RVP_OK;
}

int rvp_plane_extract(const float *x, const float *y, const float *z, size_t n, const float *m, float t, float *in_aos,
                      size_t *n_in, float *out_aos, size_t *n_out, rvp_stream_t st) {
  RVP_CHECK(m && n_in && n_out, RVP_E_ARG, "null argument");
  *n_in = *n_out = 0;
  if (n == 0) return RVP_OK;
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  Tmp c;
  RVP_TRY(c.alloc(2 * sizeof(uint32_t), s));
  RVP_TRY(plane_split_aos(x, y, z, n, m, t, in_aos, out_aos, c.as<uint32_t>(), s));
  uint32_t h[2];
  RVP_HIP(hipMemcpyAsync(h, c.p, sizeof(h), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *n_in = h[0];
  *n_out = h[1];
  return RVP_OK;
}

int rvp_ransac_plane_extract(const float *x, const float *y, const float *z, size_t n, float dist_thresh,
                             int max_iters, float collinear_thresh, float probability, float *model, int *inliers,
                             int *iterations, float *in_aos, size_t *n_in, float *out_aos, size_t *n_out,
                             rvp_stream_t st) {
  RVP_CHECK(model && inliers && n_in && n_out, RVP_E_ARG, "null output");
  *n_in = *n_out = 0;
  RVP_TRY(rvp_ransac_plane(x, y, z, n, dist_thresh, max_iters, collinear_thresh, probability, model, inliers,
                           iterations, st));
  // pipeline_export.cpp:471-476: the split runs only when a model was found
  if (*inliers <= 0) return RVP_OK;
  return rvp_plane_extract(x, y, z, n, model, dist_thresh, in_aos, n_in, out_aos, n_out, st);
}

int rvp_batch_voxel(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                    size_t n_bound, const float *leaf, float *out_aos, uint32_t *out_offsets, uint32_t *status,
                    rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && leaf && out_aos && out_offsets && status, RVP_E_ARG, "null pointer");
  RVP_CHECK(nclouds > 0 && n_bound > 0, RVP_E_ARG, "empty batch");
  RVP_CHECK(n_bound < 0xfffffff0ull, RVP_E_UNSUPPORTED, "more than 2^32-16 points");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  SegGrid G;  // only its offset clamp is used
  RVP_TRY(G.offs.alloc((nclouds + 1) * sizeof(uint32_t), s));
  RVP_TRY(seg_clamp_offsets(offsets, nclouds, n_bound, G.offs.as<uint32_t>(), s));
  const uint32_t *off = G.offs.as<uint32_t>();
  Tmp vmm, pmm, ryz, base, bits, keys, vals, kt, vt, vid, start, nv;
  const unsigned per = (unsigned)std::min<size_t>(256, std::max<size_t>(1, n_bound / ((size_t)nclouds * 2048)));
  RVP_TRY(vmm.alloc(6 * nclouds * sizeof(int), s));
  RVP_TRY(pmm.alloc((size_t)6 * nclouds * per * sizeof(int), s));
  RVP_TRY(ryz.alloc(2 * nclouds * sizeof(uint64_t), s));
  RVP_TRY(base.alloc((nclouds + 1) * sizeof(uint64_t), s));
  RVP_TRY(bits.alloc(sizeof(int), s));
  RVP_LAUNCH("seg_vox_minmax_k", s, seg_vox_minmax_k<<<dim3(per, nclouds), 256, 0, s>>>(x, y, z, off, leaf,
                                                                                     pmm.as<int>()));
  if (nclouds <= VOX_BOUNDS_MAX) {
    RVP_LAUNCH("seg_vox_bounds_k", s, seg_vox_bounds_k<<<1, 1024, 0, s>>>(pmm.as<int>(), (int)per, nclouds, off,
                                                                          vmm.as<int>(), ryz.as<uint64_t>(),
                                                                          base.as<uint64_t>(), bits.as<int>(), status));
  } else {
    RVP_LAUNCH("seg_vox_reduce_k", s, seg_vox_reduce_k<<<nclouds, 64, 0, s>>>(pmm.as<int>(), (int)per,
                                                                              vmm.as<int>()));
    RVP_LAUNCH("seg_vox_base_k", s, seg_vox_base_k<<<1, 64, 0, s>>>(nclouds, off, vmm.as<int>(), ryz.as<uint64_t>(),
                                                                    base.as<uint64_t>(), bits.as<int>(), status));
  }
  RVP_TRY(keys.alloc(n_bound * sizeof(uint64_t), s));
  RVP_TRY(vals.alloc(n_bound * sizeof(uint32_t), s));
  RVP_TRY(kt.alloc(n_bound * sizeof(uint64_t), s));
  RVP_TRY(vt.alloc(n_bound * sizeof(uint32_t), s));
  const unsigned nb = ceil_div(n_bound, 256);
  RVP_LAUNCH("seg_vox_keys_k", s, seg_vox_keys_k<<<nb, 256, 0, s>>>(x, y, z, off, nclouds, leaf, vmm.as<int>(),
                                                                    ryz.as<uint64_t>(), base.as<uint64_t>(), status,
                                                                    keys.as<uint64_t>(), vals.as<uint32_t>()));
  uint64_t *kp = keys.as<uint64_t>(), *ktp = kt.as<uint64_t>();
  uint32_t *vp = vals.as<uint32_t>(), *vtp = vt.as<uint32_t>();
  const uint32_t *n_dev = off + nclouds;
  // every input point has a key: the bound is the count, so large sorts take the large tiles.  Each cloud
  // spans fewer than 2^48 keys (seg_vox_base_k), so the keys fit 48 + bits_for(nclouds - 1) bits: the
  // passes above that are not launched at all (the device-side width skips the empty ones below it).
  const int key_bits = std::min(64, 48 + bits_for((uint64_t)nclouds - 1));
  RVP_TRY(radix_sort_pairs_u64(&kp, &vp, &ktp, &vtp, n_bound, key_bits, s, n_dev, bits.as<int>(),
                               n_bound >= (1u << 19)));
  RVP_TRY(vid.alloc(n_bound * sizeof(uint32_t), s));
  RVP_TRY(nv.alloc(sizeof(uint32_t), s));
  RVP_TRY(start.alloc(n_bound * sizeof(uint32_t), s));
  RVP_TRY(run_heads(kp, n_bound, n_dev, vid.as<uint32_t>(), start.as<uint32_t>(), nv.as<uint32_t>(), s));
  RVP_LAUNCH("voxel_centroid_k", s, voxel_centroid_k<<<nb, 256, 0, s>>>(x, y, z, vp, start.as<uint32_t>(),
                                                                        nv.as<uint32_t>(), (uint32_t)n_bound, n_dev,
                                                                        out_aos));
  RVP_LAUNCH("seg_vox_offsets_k", s, seg_vox_offsets_k<<<ceil_div(nclouds + 1, 64), 64, 0, s>>>(
                                         off, nclouds, vid.as<uint32_t>(), nv.as<uint32_t>(), out_offsets));
  return RVP_OK;
}

int rvp_batch_ransac_plane(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                           size_t n_bound, float dist_thresh, int max_iters, float collinear_thresh,
                           float probability, float *models_out, int *inliers_out, int *iterations_out,
                           rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && models_out && inliers_out, RVP_E_ARG, "null pointer");
  RVP_CHECK(nclouds > 0 && n_bound > 0, RVP_E_ARG, "empty batch");
  RVP_CHECK(n_bound < 0xfffffff0ull, RVP_E_UNSUPPORTED, "more than 2^32-16 points");
  RVP_CHECK(max_iters >= 0 && (size_t)nclouds * (size_t)std::max(max_iters, 1) < (1ull << 31), RVP_E_UNSUPPORTED,
            "clouds x hypotheses beyond 2^31");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  const int H = std::max(max_iters, 1);
  const size_t NH = (size_t)nclouds * H;
  SegGrid G;
  RVP_TRY(G.offs.alloc((nclouds + 1) * sizeof(uint32_t), s));
  RVP_TRY(seg_clamp_offsets(offsets, nclouds, n_bound, G.offs.as<uint32_t>(), s));
  const uint32_t *off = G.offs.as<uint32_t>();
  const int32_t *draws = nullptr;
  RVP_TRY(rand_draws((size_t)3 * H, s, &draws));
  // models [NH][4], valid [NH], counts [NH], offsets [nclouds + 1]: one buffer, read back in one copy
  const size_t words = NH * 6 + (size_t)nclouds + 1;
  Tmp pack, tiles;
  RVP_TRY(pack.alloc(words * sizeof(uint32_t), s));
  float *models = pack.as<float>();
  int32_t *valid = reinterpret_cast<int32_t *>(models + NH * 4), *counts = valid + NH;
  uint32_t *offs_copy = reinterpret_cast<uint32_t *>(counts + NH);
  RVP_TRY(tiles.alloc((nclouds + 1) * sizeof(uint32_t), s));
  // Hypotheses are counted in rounds of 64, 128, ... and replayed after each: the reference's loop usually
  // stops long before max_iters (adaptively, ransac_plane.cpp:95-105), and a cloud whose loop has stopped
  // takes no part in later rounds.  The first read-back brings the models, flags and offsets too.  The models
  // launch also lays out the count tiles and marks every cloud active for the first round.
  Tmp act;
  RVP_TRY(act.alloc((size_t)nclouds, s));
  RVP_LAUNCH("seg_plane_models_k", s, seg_plane_models_k<<<ceil_div(NH, 256), 256, 0, s>>>(
                                          x, y, z, off, nclouds, draws, H, collinear_thresh, models, valid, counts,
                                          tiles.as<uint32_t>(), offs_copy, act.as<uint8_t>()));
  PinnedStage stage;
  uint8_t *host = nullptr;
  const size_t act_at = (words * sizeof(uint32_t) + 255) & ~(size_t)255;
  RVP_TRY(stage.get(s, act_at + (size_t)nclouds, reinterpret_cast<void **>(&host)));
  uint8_t *hact = host + act_at;
  const float *hm = reinterpret_cast<const float *>(host);
  const int32_t *hv = reinterpret_cast<const int32_t *>(hm + NH * 4), *hc = hv + NH;
  const uint32_t *ho = reinterpret_cast<const uint32_t *>(hc + NH);
  std::vector<RansacReplay> rps;
  int h_begin = 0, chunk = 64;
  for (;;) {
    const int h_end = std::min(H, h_begin + chunk);
    const dim3 grid((unsigned)(ceil_div(n_bound, RC_TILE) + nclouds), ceil_div((size_t)(h_end - h_begin), RC_HYP));
    RVP_LAUNCH("seg_plane_count_k", s, seg_plane_count_k<<<grid, 256, 0, s>>>(
                                           x, y, z, off, nclouds, tiles.as<uint32_t>(), models, valid, H, h_begin,
                                           h_end, act.as<uint8_t>(), dist_thresh, counts));
    if (h_begin == 0) RVP_HIP(hipMemcpyAsync(host, pack.p, words * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    else RVP_HIP(hipMemcpyAsync(host + NH * 5 * sizeof(int32_t), counts, NH * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    RVP_HIP(hipStreamSynchronize(s));
    if (h_begin == 0) {
      rps.reserve(nclouds);
      for (int c = 0; c < nclouds; ++c) {
        const size_t n = ho[c + 1] - ho[c];
        rps.emplace_back(n, max_iters, probability);
        if (n < 3) rps.back().done = true;  // ransac_plane.cpp:59
      }
    }
    bool more = false;
    for (int c = 0; c < nclouds; ++c) {
      RansacReplay &rp = rps[c];
      if (!rp.done) {
        const size_t b = (size_t)c * H + h_begin;
        rp.feed(hm + b * 4, hv + b, hc + b, h_end - h_begin);
      }
      hact[c] = rp.done ? 0 : 1;
      more = more || !rp.done;
    }
    if (!more || h_end >= H) break;
    RVP_HIP(hipMemcpyAsync(act.p, hact, (size_t)nclouds, hipMemcpyHostToDevice, s));
    h_begin = h_end;
    chunk = std::min(chunk * 2, 1024);
  }
  for (int c = 0; c < nclouds; ++c) {
    const size_t n = ho[c + 1] - ho[c];
    for (int k = 0; k < 4; ++k) models_out[4 * c + k] = 0.0f;
    inliers_out[c] = 0;
    if (iterations_out) iterations_out[c] = 0;
    if (n < 3) continue;
    const RansacReplay &rp = rps[c];
    for (int k = 0; k < 4; ++k) models_out[4 * c + k] = rp.model[k];
    inliers_out[c] = rp.best;
    if (iterations_out) iterations_out[c] = rp.iterations();
  }
  return RVP_OK;
}

}  // extern "C"
