// normals.hip -- N1 / N2 normal estimation: neighbourhood (exact kNN or radius list), covariance re-centred
// at the query point with fp64 terms and sums, closed-form FP32 3x3 eigen solve with an FP64 Jacobi
// fallback for near-degenerate spectra, flip toward the viewpoint.
//
// Two kernels per call: the neighbourhood kernels (one wave per query) write each query's moments to a
// record array, and normals_solve_k solves the eigen problems one lane per query.  Keeping the solver out of
// the search kernels keeps them near 50-60 VGPRs (seven waves per SIMD instead of four or five), and the
// solver's instructions serve 64 queries per wave instead of the handful a search wave holds.
//
// Reference: normal_estimation_sc (src/features/normal_estimation.cpp:72-112), normal_estimation_rvv with
// a SpatialHash (:189-214), flipNormalTowardsViewpoint (:55-67).  The reference's own eigen solver
// (simple_eigen3x3_smallest, :16-52: four non-converged Jacobi steps with a sign error) is deliberately not
// reproduced; results are graded against an FP64 eigen solve of the same neighbourhood.
#include <map>
#include <mutex>
#include <type_traits>

#include "batch.hpp"
#include "prims.hpp"
#include "query.hpp"

namespace rvp {

// FP64-fallback counters: one device word per (device, stream), allocated once under a lock and kept until the
// process ends (a stream handle the driver reuses gets the same word back).  A normals call's neighbourhood kernel
// clears its stream's word (block 0, in stream order: no memset node, so the call can be captured into a graph)
// and its solve kernel adds to it; calls on other streams never touch it.  Each host thread remembers the stream of its last
// normals call, which rvp_normals_fallback_count reads.
static std::mutex g_fallback_mu;
static std::map<std::pair<int, hipStream_t>, unsigned long long *> g_fallback;
struct LastNormalsCall {
  int dev = -1;
  hipStream_t s = nullptr;
  unsigned long long *counter = nullptr;
};
static thread_local LastNormalsCall g_last_normals;

static int fallback_counter(hipStream_t s, unsigned long long **out) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  unsigned long long *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_fallback_mu);
    unsigned long long *&slot = g_fallback[{dev, s}];
    if (!slot) RVP_HIP(hipMalloc(&slot, sizeof(unsigned long long)));
    p = slot;
  }
  g_last_normals = LastNormalsCall{dev, s, p};
  *out = p;
  return RVP_OK;
}

// Each wave handles NB consecutive queries (cell order): from the second on, a k-NN query is bounded by the
// previous one's K-th distance, and radius windows stage once in LDS for the queries that share them.
constexpr int NB = 8;

// Per-query record of the neighbourhood kernels, structure of arrays with stride n (cell-order position q):
// rec[e * n + q] for e = 0..5 the covariance (xx, xy, xz, yy, yz, zz), 6..8 the neighbourhood mean relative
// to the query point, 9 the neighbour count (int bits).
constexpr int REC_WORDS = 10;

// cyclic Jacobi in double precision (symmetric 3x3), eigenvector of the smallest eigenvalue
__device__ void jacobi_smallest_f64(double a00, double a01, double a02, double a11, double a12, double a22,
                                    double v[3]) {
  // every index below is a compile-time constant (unrolled loops, selects for the result column), so the
  // arrays live in registers and the kernels need no scratch memory
  double A[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
  double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    const double dia = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
    if (off == 0.0 || off < 1e-18 * dia) break;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = p + 1; q < 3; ++q) {
        const double apq = A[p][q];
        if (apq != 0.0) {
          const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
          const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double akp = A[k][p], akq = A[k][q];
            A[k][p] = c * akp - s * akq;
            A[k][q] = s * akp + c * akq;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double apk = A[p][k], aqk = A[q][k];
            A[p][k] = c * apk - s * aqk;
            A[q][k] = s * apk + c * aqk;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double vkp = V[k][p], vkq = V[k][q];
            V[k][p] = c * vkp - s * vkq;
            V[k][q] = s * vkp + c * vkq;
          }
        }
      }
  }
  // the column of the smallest diagonal entry (the first one on a tie)
  const bool c1 = A[1][1] < A[0][0];
  const double dmin = c1 ? A[1][1] : A[0][0];
  const bool c2 = A[2][2] < dmin;
  double u[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) u[k] = c2 ? V[k][2] : (c1 ? V[k][1] : V[k][0]);
  const double nn = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  v[0] = u[0] / nn;
  v[1] = u[1] / nn;
  v[2] = u[2] / nn;
}

// Smallest-eigenvalue eigenvector of a symmetric 3x3 (Smith's trigonometric eigenvalues, eigenvector from
// the largest cross product of two rows of A - l3 I).  Returns true when it fell back to FP64 Jacobi: a
// relative gap (l2 - l3) / l1 that may lie below 1e-4 once the rounding of the trigonometric solve is
// allowed for, or an isotropic spectrum.
__device__ bool smallest_eigvec(const float c[6], float n[3]) {
  float a00 = c[0], a01 = c[1], a02 = c[2], a11 = c[3], a12 = c[4], a22 = c[5];
  float scale = fmaxf(fmaxf(fmaxf(fabsf(a00), fabsf(a01)), fmaxf(fabsf(a02), fabsf(a11))),
                      fmaxf(fabsf(a12), fabsf(a22)));
  if (!(scale > 0.0f) || !(scale < FLT_MAX)) {
    n[0] = 1.0f; n[1] = 0.0f; n[2] = 0.0f;
    return false;
  }
  const float is = 1.0f / scale;
  a00 *= is; a01 *= is; a02 *= is; a11 *= is; a12 *= is; a22 *= is;
  const float q = (a00 + a11 + a22) * (1.0f / 3.0f);
  const float b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const float p2 = (b00 * b00 + b11 * b11 + b22 * b22 + 2.0f * (a01 * a01 + a02 * a02 + a12 * a12)) * (1.0f / 6.0f);
  bool fallback = !(p2 > 1e-30f);
  if (!fallback) {
    const float p = sqrtf(p2), ip = 1.0f / p;
    const float B00 = b00 * ip, B11 = b11 * ip, B22 = b22 * ip, B01 = a01 * ip, B02 = a02 * ip, B12 = a12 * ip;
    const float det = B00 * (B11 * B22 - B12 * B12) - B01 * (B01 * B22 - B12 * B02) + B02 * (B01 * B12 - B11 * B02);
    const float r = fminf(fmaxf(0.5f * det, -1.0f), 1.0f);
    const float phi = acosf(r) * (1.0f / 3.0f);
    const float l1 = q + 2.0f * p * cosf(phi);
    const float l3 = q + 2.0f * p * cosf(phi + 2.0943951023931953f);
    const float l2 = 3.0f * q - l1 - l3;
    // Near a double smallest root (phi -> 0) acos loses half the digits of r: with r within 2^-18 of its true
    // value, phi is off by up to sqrt(2 * 2^-18) / 3 < 9.2e-4, so l2 - l3 = 2 sqrt(3) p sin(phi) is off by up to
    // 3.2e-3 p.  The FP64 solve therefore also takes every spectrum whose computed gap does not exceed the
    // 1e-4 relative threshold by that bound: a true gap below 1e-4 |l1| never reaches the closed form.
    if (!((l2 - l3) >= 1e-4f * fmaxf(fabsf(l1), 1e-30f) + 3.2e-3f * p)) {
      fallback = true;
    } else {
      const float r0[3] = {a00 - l3, a01, a02}, r1[3] = {a01, a11 - l3, a12}, r2[3] = {a02, a12, a22 - l3};
      float v[3][3];
      v[0][0] = r0[1] * r1[2] - r0[2] * r1[1]; v[0][1] = r0[2] * r1[0] - r0[0] * r1[2]; v[0][2] = r0[0] * r1[1] - r0[1] * r1[0];
      v[1][0] = r0[1] * r2[2] - r0[2] * r2[1]; v[1][1] = r0[2] * r2[0] - r0[0] * r2[2]; v[1][2] = r0[0] * r2[1] - r0[1] * r2[0];
      v[2][0] = r1[1] * r2[2] - r1[2] * r2[1]; v[2][1] = r1[2] * r2[0] - r1[0] * r2[2]; v[2][2] = r1[0] * r2[1] - r1[1] * r2[0];
      // the longest cross product (the first one on a tie), chosen by selects so that v stays in registers
      float best = -1.0f, b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float nn = v[i][0] * v[i][0] + v[i][1] * v[i][1] + v[i][2] * v[i][2];
        if (nn > best) { best = nn; b0 = v[i][0]; b1 = v[i][1]; b2 = v[i][2]; }
      }
      if (!(best > 0.0f)) {
        fallback = true;
      } else {
        const float inv = 1.0f / sqrtf(best);
        n[0] = b0 * inv; n[1] = b1 * inv; n[2] = b2 * inv;
      }
    }
  }
  if (fallback) {
    double v[3];
    jacobi_smallest_f64(c[0], c[1], c[2], c[3], c[4], c[5], v);
    n[0] = (float)v[0]; n[1] = (float)v[1]; n[2] = (float)v[2];
  }
  return fallback;
}

// flipNormalTowardsViewpoint (normal_estimation.cpp:55-67), fp32 in the reference's order
__device__ __forceinline__ void flip_toward(float px, float py, float pz, float vpx, float vpy, float vpz, float n[3]) {
  const float vx = __fsub_rn(vpx, px), vy = __fsub_rn(vpy, py), vz = __fsub_rn(vpz, pz);
  const float dot = __fadd_rn(__fadd_rn(__fmul_rn(vx, n[0]), __fmul_rn(vy, n[1])), __fmul_rn(vz, n[2]));
  if (dot < 0.0f) {
    n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2];
  }
}

// Covariance of the m <= 64 neighbours held one per lane, coordinates re-centred at the query point, with
// sums that do not depend on which lane holds which neighbour (a neighbourhood found without sorting arrives
// in an order that can vary between runs).  Every lane rounds its offsets p, and its products p_a p_b, to
// integer multiples of 2^(E - 46), resp. 2^(2E - 46), where 2^E bounds every |p| of the wave: at most 64
// terms below 2^46 units each sum exactly in fp64, so the totals are the same for any lane order.  The
// rounding is 2^-47 of the bound, far below fp32.  Then sum((p - mu)(p - mu)^T) = sum(p p^T) - m mu mu^T.
// Result uniform across the wave, rounded to fp32; c[6..8] receives the neighbourhood mean relative to the
// query.
__device__ __forceinline__ void wave_cov_unordered(bool valid, float px, float py, float pz, int m, float qx,
                                                   float qy, float qz, float c[9]) {
  double p[3] = {0.0, 0.0, 0.0};
  if (valid) {
    p[0] = (double)px - (double)qx;
    p[1] = (double)py - (double)qy;
    p[2] = (double)pz - (double)qz;
  }
  // E: the wave's largest binary exponent of |p| (|p| < 2^E)
  int e = -1074;
  const double am = fmax(fmax(fabs(p[0]), fabs(p[1])), fabs(p[2]));
  if (am > 0.0) e = __builtin_amdgcn_frexp_exp(am);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) e = max(e, shfl_xor(e, d));
  e = max(e, -460);  // all-zero offsets (or nearly): any scale leaves them exact
  double s1[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) s1[a] = rint(ldexp(p[a], 46 - e));
  const double pr[6] = {rint(ldexp(p[0] * p[0], 46 - 2 * e)), rint(ldexp(p[0] * p[1], 46 - 2 * e)),
                        rint(ldexp(p[0] * p[2], 46 - 2 * e)), rint(ldexp(p[1] * p[1], 46 - 2 * e)),
                        rint(ldexp(p[1] * p[2], 46 - 2 * e)), rint(ldexp(p[2] * p[2], 46 - 2 * e))};
  double mu[3], s2[6];
  wave_sum3_f64(s1, mu);
  wave_sum6_f64(pr, s2);
  const double inv = m > 0 ? 1.0 / (double)m : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) mu[a] = ldexp(mu[a], e - 46) * inv;
  const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int k = 0; k < 6; ++k) c[k] = (float)(ldexp(s2[k], 2 * e - 46) - (double)m * mu[ia[k]] * mu[ib[k]]);
#pragma unroll
  for (int a = 0; a < 3; ++a) c[6 + a] = (float)mu[a];
}

// The same for neighbours held in a canonical lane order (sorted by (d2, idx)), whose fp64 sums are therefore
// reproducible as they stand: two passes (mean, then centred products) with fp64 terms and sums.  UNORDERED
// selects wave_cov_unordered (the same results for any lane order).  The neighbour is at cell-order position
// pos of the grid.
template <bool UNORDERED = false>
__device__ __forceinline__ void wave_cov(const GridView &g, bool valid, uint32_t pos, int m, float qx, float qy,
                                         float qz, float c[9]) {
  float px = 0.0f, py = 0.0f, pz = 0.0f;
  if (valid) {
    px = g.xs[pos];
    py = g.ys[pos];
    pz = g.zs[pos];
  }
  if (UNORDERED) {
    wave_cov_unordered(valid, px, py, pz, m, qx, qy, qz, c);
    return;
  }
  double p[3] = {0.0, 0.0, 0.0};
  if (valid) {
    p[0] = (double)px - (double)qx;
    p[1] = (double)py - (double)qy;
    p[2] = (double)pz - (double)qz;
  }
  double mu[3];
  wave_sum3_f64(p, mu);
  const double inv = m > 0 ? 1.0 / (double)m : 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) mu[a] *= inv;
  const double dx = valid ? p[0] - mu[0] : 0.0, dy = valid ? p[1] - mu[1] : 0.0, dz = valid ? p[2] - mu[2] : 0.0;
  const double pr[6] = {dx * dx, dx * dy, dx * dz, dy * dy, dy * dz, dz * dz};
  double cv[6];
  wave_sum6_f64(pr, cv);
#pragma unroll
  for (int e = 0; e < 6; ++e) c[e] = (float)cv[e];
#pragma unroll
  for (int a = 0; a < 3; ++a) c[6 + a] = (float)mu[a];
}

// the wave's moments of query q (uniform across the wave after wave_cov) into the record: lane e stores word e
__device__ __forceinline__ void write_rec(float *__restrict__ rec, size_t n, uint32_t q, const float c[9], int m) {
  const int l = lane();
  float v = __int_as_float(m);
#pragma unroll
  for (int e = 0; e < 9; ++e)
    if (l == e) v = c[e];
  if (l < REC_WORDS) rec[(size_t)l * n + q] = v;
}

// One lane per query: eigen solve of the recorded covariance (m >= 3 neighbours, else a zero normal), flip
// toward the viewpoint, scatter to the query's row (perm: cell order -> caller's order).  count_dev: the
// device-side query count (batched stages), else n.  cx may be null; else cx/cy/cz receive the neighbourhood
// centroid q + mean.
__global__ __launch_bounds__(256) void normals_solve_k(const float *__restrict__ rec, size_t n,
                                                        const uint32_t *__restrict__ count_dev,
                                                        const float *__restrict__ xs, const float *__restrict__ ys,
                                                        const float *__restrict__ zs, const uint32_t *__restrict__ perm,
                                                        float vpx, float vpy, float vpz, float *__restrict__ nx,
                                                        float *__restrict__ ny, float *__restrict__ nz,
                                                        unsigned long long *fallbacks, float *__restrict__ cx,
                                                        float *__restrict__ cy, float *__restrict__ cz) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t count = count_dev ? min((size_t)*count_dev, n) : n;
  bool fb = false;
  if (q < count) {
    const int m = __float_as_int(rec[9 * n + q]);
    const float px = xs[q], py = ys[q], pz = zs[q];
    float nrm[3] = {0.f, 0.f, 0.f};
    if (m >= 3) {
      const float c[6] = {rec[q], rec[n + q], rec[2 * n + q], rec[3 * n + q], rec[4 * n + q], rec[5 * n + q]};
      fb = smallest_eigvec(c, nrm);
      flip_toward(px, py, pz, vpx, vpy, vpz, nrm);
    }
    const uint32_t r = perm[q];
    nx[r] = nrm[0];
    ny[r] = nrm[1];
    nz[r] = nrm[2];
    if (cx) {
      cx[r] = __fadd_rn(px, rec[6 * n + q]);
      cy[r] = __fadd_rn(py, rec[7 * n + q]);
      cz[r] = __fadd_rn(pz, rec[8 * n + q]);
    }
  }
  const int nf = popc(ballot(fb));
  if (lane() == 0 && nf) atomicAdd(fallbacks, (unsigned long long)nf);
}

// MODE 0: k nearest (N1, K = k+1 incl. self); MODE 1: radius list capped at K (N2)
template <int MODE, bool STRIDED>
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(MODE == 1 ? 7 : 8)
void normals_topk_k(GridView g, int K, float radius, int h0, int nb, float *__restrict__ rec,
                    unsigned long long *clear) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0ull;  // the call's fallback counter
  __shared__ WaveLds lds[QW];
  // consecutive queries of a batch (not strided) mostly share their window: stage it once in LDS
  constexpr bool STAGE = MODE == 1 && !STRIDED;
  __shared__ typename std::conditional<STAGE, StageLds, char>::type stage[QW];
  const int w = wave_in_block();
  WaveLds &L = lds[w];
  const float r2 = __fmul_rn(radius, radius);
  uint32_t first = 0, step = 1;
  int nbatch = 0;
  for (uint32_t k = 0; wave_batch<STRIDED>(k, g.n, nb, first, step, nbatch); ++k) {
    StagedWindow win;
    float px = 0.f, py = 0.f, pz = 0.f, pd2 = -1.0f;  // previous query of the batch, its K-th d2 (< 0: none)
    for (int j = 0; j < nbatch; ++j) {
      const uint32_t q = first + (uint32_t)j * step;
      const float x = uniform(g.xs[q]), y = uniform(g.ys[q]), z = uniform(g.zs[q]);
      TopK t;
      t.init(K);
      if (MODE == 0) {
        // consecutive queries: the previous one's K-th distance bounds this one's (knn_bound_from)
        knn_search<false>(g, x, y, z, t, L, STRIDED ? __builtin_inff() : knn_bound_from(x, y, z, px, py, pz, pd2),
                          h0);
        px = x;
        py = y;
        pz = z;
        pd2 = t.cnt >= t.K ? kd2(t.kth) : -1.0f;
      } else {
        int lo[3], hi[3];
        window3(g, x, y, z, radius, lo, hi);
        TopKSink<false> sink{t, L};
        if constexpr (STAGE) {
          win.ensure(g, lo, hi, 0, L, stage[w]);
          if (win.total >= 0) {
            // at most K hits: they are the capped list, covariance straight from the staged coordinates
            float hx, hy, hz;
            const int h = staged_hits_upto(stage[w], win.total, x, y, z, r2, K, L, hx, hy, hz);
            if (h >= 0) {
              float c[9];
              wave_cov_unordered(lane() < h, hx, hy, hz, h, x, y, z, c);
              write_rec(rec, g.n, q, c, h);
              continue;
            }
            scan_staged(stage[w], win.total, x, y, z, r2, sink);
          } else {
            scan_box(g, x, y, z, lo, hi, false, lo, hi, r2, true, L, sink);
          }
        } else {
          scan_box(g, x, y, z, lo, hi, false, lo, hi, r2, true, L, sink);
        }
        topk_flush<false>(t, L);
      }
      const int m = min(t.cnt, K);
      float c[9];
      // the list holds (d2, index) keys only: the neighbour's cell-order position is rank[index]; radius lists
      // use the formula of their unordered fast path
      wave_cov<MODE == 1>(g, lane() < m, lane() < m ? g.rank[key_idx(t.key)] : 0u, m, x, y, z, c);
      write_rec(rec, g.n, q, c, m);
    }
  }
}

// N2 without a cap: neighbour positions in CSR rows (row = original index), covariance in chunks of 64
__global__ __launch_bounds__(QT) void normals_csr_k(GridView g, const int64_t *__restrict__ offsets, int cap,
                                                     const uint32_t *__restrict__ pos, float *__restrict__ rec,
                                                     unsigned long long *clear) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0ull;  // the call's fallback counter
  const int w = wave_in_block();
  const uint32_t base = (blockIdx.x * QW + w) * NB;
  if (base >= g.n) return;
  const int l = lane();
  const int nbatch = (int)min<uint32_t>(NB, g.n - base);
  for (int j = 0; j < nbatch; ++j) {
    const uint32_t q = base + j;
    const uint32_t row = g.perm[q];
    const float x = g.xs[q], y = g.ys[q], z = g.zs[q];
    const int64_t b = offsets[row];
    const int m = (int)min<int64_t>(offsets[row + 1] - b, (int64_t)cap);
    float c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (m <= WAVE) {
      wave_cov(g, l < m, l < m ? pos[b + l] : 0u, m, x, y, z, c);
    } else {
      // two passes over the row; each lane keeps fp64 running sums across chunks
      double sp[3] = {0.0, 0.0, 0.0};
      for (int c0 = 0; c0 < m; c0 += WAVE) {
        if (c0 + l < m) {
          const uint32_t p = pos[b + c0 + l];
          sp[0] += (double)g.xs[p] - (double)x;
          sp[1] += (double)g.ys[p] - (double)y;
          sp[2] += (double)g.zs[p] - (double)z;
        }
      }
      double mu[3];
      wave_sum3_f64(sp, mu);
      const double inv = 1.0 / (double)m;
#pragma unroll
      for (int a = 0; a < 3; ++a) mu[a] *= inv;
      double cs[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int c0 = 0; c0 < m; c0 += WAVE) {
        if (c0 + l < m) {
          const uint32_t p = pos[b + c0 + l];
          const double d[3] = {(double)g.xs[p] - (double)x - mu[0], (double)g.ys[p] - (double)y - mu[1],
                               (double)g.zs[p] - (double)z - mu[2]};
          cs[0] += d[0] * d[0];
          cs[1] += d[0] * d[1];
          cs[2] += d[0] * d[2];
          cs[3] += d[1] * d[1];
          cs[4] += d[1] * d[2];
          cs[5] += d[2] * d[2];
        }
      }
      double cv[6];
      wave_sum6_f64(cs, cv);
#pragma unroll
      for (int e = 0; e < 6; ++e) c[e] = (float)cv[e];
#pragma unroll
      for (int a = 0; a < 3; ++a) c[6 + a] = (float)mu[a];
    }
    write_rec(rec, g.n, q, c, m);
  }
}

// Capped radius lists of the same clouds (rvp_batch_radius_search's rows: stride cap, indexed by batch position,
// neighbour indices local to the cloud, (d2, idx) order), read by the k-NN normals in place of a search.
struct NbrLists {
  const int32_t *idx = nullptr;
  const float *d2 = nullptr;  // may be null
  const int32_t *cnt = nullptr;
  int cap = 0;
};

// N1 / N2 over a batch: every query's neighbourhood within its own cloud (MODE 0: K nearest incl. self;
// MODE 1: radius[c] list capped at K).  LISTS (MODE 0, K <= lists.cap): a query whose radius list holds at
// least K entries takes them as its K nearest: every point outside the radius has a larger d2 than every
// point inside, and both orders are (d2, idx) on the same coordinates, so the first K entries of the list
// are the K nearest.  Lane j then holds entry j, as after a search, and the moments are the searched ones.
template <int MODE, bool STRIDED, bool LISTS = false>
__global__ __launch_bounds__(QT) RVP_WAVES_PER_SIMD(MODE == 1 ? 7 : 8)
void seg_normals_k(SegArrays A, int K, const float *__restrict__ radius, int h0, int nb, float *__restrict__ rec,
                   size_t n_bound, unsigned long long *clear, NbrLists lists) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *clear = 0ull;  // the call's fallback counter
  __shared__ WaveLds lds[QW];
  // consecutive queries of a batch (not strided) mostly share their window: stage it once in LDS
  constexpr bool STAGE = MODE == 1 && !STRIDED;
  __shared__ typename std::conditional<STAGE, StageLds, char>::type stage[QW];
  const int w = wave_in_block();
  const uint32_t total = A.off[A.ncl];
  uint32_t first = 0, step = 1;
  int nbatch = 0;
  for (uint32_t k = 0; wave_batch<STRIDED>(k, total, nb, first, step, nbatch); ++k) {
    StagedWindow win;
    int pc = -1;  // cloud of the previous query of the batch, whose K-th d2 is pd2
    float px = 0.f, py = 0.f, pz = 0.f, pd2 = -1.0f;
    for (int j = 0; j < nbatch; ++j) {
      const uint32_t q = first + (uint32_t)j * step;
      const int c = cloud_of(A.off, A.ncl, q);
      const GridView g = seg_view(A, c);
      const float x = uniform(g.xs[q]), y = uniform(g.ys[q]), z = uniform(g.zs[q]);
      TopK t;
      t.init(K);
      if (MODE == 0) {
        bool listed = false;
        if constexpr (LISTS) {
          const uint32_t row = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.perm[q]);
          if (lists.cnt[row] >= K) {
            const uint32_t base = A.off[c], size = A.off[c + 1] - base;
            const size_t at = (size_t)row * (uint32_t)lists.cap;
            const int32_t li = lane() < K ? lists.idx[at + lane()] : 0;
            // an index outside the cloud (a list not made for these clouds) falls back to the search
            if (!ballot(lane() < K && (li < 0 || (uint32_t)li >= size))) {
              t.key = lane() < K ? pack_key(0.0f, base + (uint32_t)li) : ~0ull;
              t.cnt = K;
              listed = true;
              pd2 = lists.d2 ? uniform(lists.d2[at + K - 1]) : -1.0f;
            }
          }
        }
        if (!listed) {
          // consecutive queries of one cloud: the previous one's K-th distance bounds this one's
          knn_search<false>(g, x, y, z, t, lds[w],
                            (!STRIDED && c == pc) ? knn_bound_from(x, y, z, px, py, pz, pd2) : __builtin_inff(),
                            h0);
          pd2 = t.cnt >= t.K ? kd2(t.kth) : -1.0f;
        }
        pc = c;
        px = x;
        py = y;
        pz = z;
      } else {
        const float r = radius[c];
        const float r2 = __fmul_rn(r, r);
        int lo[3], hi[3];
        window3(g, x, y, z, r, lo, hi);
        TopKSink<false> sink{t, lds[w]};
        if constexpr (STAGE) {
          win.ensure(g, lo, hi, c, lds[w], stage[w]);
          if (win.total >= 0) {
            // at most K hits: they are the capped list, covariance straight from the staged coordinates
            float hx, hy, hz;
            const int h = staged_hits_upto(stage[w], win.total, x, y, z, r2, K, lds[w], hx, hy, hz);
            if (h >= 0) {
              float c6[9];
              wave_cov_unordered(lane() < h, hx, hy, hz, h, x, y, z, c6);
              write_rec(rec, n_bound, q, c6, h);
              continue;
            }
            scan_staged(stage[w], win.total, x, y, z, r2, sink);
          } else {
            scan_box(g, x, y, z, lo, hi, false, lo, hi, r2, true, lds[w], sink);
          }
        } else {
          scan_box(g, x, y, z, lo, hi, false, lo, hi, r2, true, lds[w], sink);
        }
        topk_flush<false>(t, lds[w]);
      }
      const int m = min(t.cnt, K);
      float c6[9];
      wave_cov<MODE == 1>(g, lane() < m, lane() < m ? g.rank[key_idx(t.key)] : 0u, m, x, y, z, c6);
      write_rec(rec, n_bound, q, c6, m);
    }
  }
}

// consecutive queries per wave of the radius normals (a batch shares its staged windows): the normals_per_wave
// knob, else NB.  Fewer per wave balance small counts better (the launch's last round of waves runs part full).
static int radius_normals_batch() {
  const int forced = tune_normals_per_wave();
  return forced > 0 ? std::min(forced, WAVE) : NB;
}

// the eigen solves of n recorded queries (count_dev: their device-side count, or null for n)
static int solve_normals(const float *rec, size_t n, const uint32_t *count_dev, const float *xs, const float *ys,
                         const float *zs, const uint32_t *perm, float vpx, float vpy, float vpz, float *nx, float *ny,
                         float *nz, unsigned long long *fb, float *cx, float *cy, float *cz, hipStream_t s) {
  RVP_LAUNCH("normals_solve_k", s, normals_solve_k<<<ceil_div(n, 256), 256, 0, s>>>(
                                       rec, n, count_dev, xs, ys, zs, perm, vpx, vpy, vpz, nx, ny, nz, fb, cx, cy, cz));
  return RVP_OK;
}

}  // namespace rvp

using namespace rvp;


extern "C" {

int rvp_normals_knn_centroids(const float *x, const float *y, const float *z, size_t n, int k, float vpx, float vpy,
                              float vpz, float *nx, float *ny, float *nz, float *cx, float *cy, float *cz,
                              rvp_stream_t st) {
  RVP_CHECK((cx && cy && cz) || (!cx && !cy && !cz), RVP_E_ARG, "centroid outputs: all or none");
  RVP_CHECK(k >= 1 && k + 1 <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 63]");
  RVP_CHECK(nx && ny && nz, RVP_E_ARG, "null output");
  if (n == 0) return RVP_OK;
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  Grid G;
  RVP_TRY(G.build_for_knn(x, y, z, n, k + 1, s));
  unsigned long long *fb = nullptr;
  RVP_TRY(fallback_counter(s, &fb));
  Tmp rec;
  RVP_TRY(rec.alloc((size_t)REC_WORDS * n * sizeof(float), s));
  unsigned qb = 1;
  bool strided = false;
  RVP_TRY(batch_blocks<NB>(normals_topk_k<0, true>, n, &qb, &strided));
  const GridView g = G.view();
  const int h0 = tune_knn_start_cells();
  if (strided)
    RVP_LAUNCH("normals_topk_k", s, normals_topk_k<0, true><<<qb, QT, 0, s>>>(g, k + 1, 0.0f, h0, NB,
                                                                              rec.as<float>(), fb));
  else
    RVP_LAUNCH("normals_topk_k", s, normals_topk_k<0, false><<<qb, QT, 0, s>>>(g, k + 1, 0.0f, h0, NB,
                                                                               rec.as<float>(), fb));
  return solve_normals(rec.as<float>(), n, nullptr, g.xs, g.ys, g.zs, g.perm, vpx, vpy, vpz, nx, ny, nz, fb, cx, cy,
                       cz, s);
}

int rvp_normals_knn(const float *x, const float *y, const float *z, size_t n, int k, float vpx, float vpy, float vpz,
                    float *nx, float *ny, float *nz, rvp_stream_t st) {
  return rvp_normals_knn_centroids(x, y, z, n, k, vpx, vpy, vpz, nx, ny, nz, nullptr, nullptr, nullptr, st);
}

int rvp_normals_radius(const rvp_hash *h, float radius, int max_nn, float vpx, float vpy, float vpz, float *nx,
                       float *ny, float *nz, rvp_stream_t st) {
  RVP_CHECK(h && h->g.built, RVP_E_STATE, "hash not built");
  RVP_CHECK(nx && ny && nz, RVP_E_ARG, "null output");
  RVP_CHECK(max_nn >= 0, RVP_E_ARG, "max_nn must be >= 0");
  const size_t n = h->g.n;
  if (n == 0) return RVP_OK;
  hipStream_t s = S(st);
  unsigned long long *fb = nullptr;
  RVP_TRY(fallback_counter(s, &fb));
  const GridView g = h->g.view();
  Tmp rec;
  RVP_TRY(rec.alloc((size_t)REC_WORDS * n * sizeof(float), s));
  if (max_nn > 0 && max_nn <= MAX_TOPK) {
    // batches of consecutive queries with staged windows (as the batched radius normals)
    const int nb = radius_normals_batch();
    const unsigned qb = ceil_div(ceil_div(n, (size_t)nb), QW);
    RVP_LAUNCH("normals_topk_k", s, normals_topk_k<1, false><<<qb, QT, 0, s>>>(g, max_nn, radius, 1, nb,
                                                                               rec.as<float>(), fb));
  } else {
    // unlimited (or a cap above 64): full lists, sorted by (d2, idx), truncated by the row offsets
    Tmp cnt, off, keys, pos;
    RVP_TRY(cnt.alloc(n * sizeof(int32_t), s));
    RVP_TRY(off.alloc((n + 1) * sizeof(int64_t), s));
    RVP_TRY(rvp_radius_count(h, nullptr, nullptr, nullptr, n, radius, cnt.as<int32_t>(), st));
    int64_t tot = 0;
    RVP_TRY(rvp_exclusive_scan(cnt.as<int32_t>(), n, off.as<int64_t>(), &tot, st));
    RVP_TRY(keys.alloc((size_t)tot * sizeof(uint64_t), s));
    RVP_TRY(pos.alloc((size_t)tot * sizeof(uint32_t), s));
    RVP_TRY(radius_csr(h->g, nullptr, nullptr, nullptr, n, radius, off.as<int64_t>(), keys.as<uint64_t>(),
                       pos.as<uint32_t>(), s));
    // a cap keeps the first min(len, max_nn) entries of each (d2, idx)-sorted row
    const int cap = max_nn > 0 ? max_nn : 0x7fffffff;
    RVP_LAUNCH("normals_csr_k", s, normals_csr_k<<<ceil_div(n, QW * NB), QT, 0, s>>>(g, off.as<int64_t>(), cap,
                                                                                     pos.as<uint32_t>(),
                                                                                     rec.as<float>(), fb));
  }
  return solve_normals(rec.as<float>(), n, nullptr, g.xs, g.ys, g.zs, g.perm, vpx, vpy, vpz, nx, ny, nz, fb, nullptr,
                       nullptr, nullptr, s);
}

static int batch_normals_knn_impl(const float *x, const float *y, const float *z, const uint32_t *offsets,
                                  int nclouds, size_t n_bound, int k, const NbrLists &lists, float vpx, float vpy,
                                  float vpz, float *nx, float *ny, float *nz, hipStream_t s) {
  RVP_TRY(pool_init());
  SegGrid G;
  RVP_TRY(G.build(x, y, z, offsets, nclouds, n_bound, true, k + 1, nullptr, s));
  unsigned long long *fb = nullptr;
  RVP_TRY(fallback_counter(s, &fb));
  Tmp rec;
  RVP_TRY(rec.alloc((size_t)REC_WORDS * n_bound * sizeof(float), s));
  unsigned qb = 1;
  bool strided = false;
  const SegArrays A = G.arrays();
  const int h0 = tune_knn_start_cells();
  if (lists.idx) {
    RVP_TRY(batch_blocks<NB>(seg_normals_k<0, true, true>, n_bound, &qb, &strided));
    if (strided)
      RVP_LAUNCH("seg_normals_k", s, seg_normals_k<0, true, true><<<qb, QT, 0, s>>>(A, k + 1, nullptr, h0, NB,
                                                                                  rec.as<float>(), n_bound, fb, lists));
    else
      RVP_LAUNCH("seg_normals_k", s, seg_normals_k<0, false, true><<<qb, QT, 0, s>>>(A, k + 1, nullptr, h0, NB,
                                                                                   rec.as<float>(), n_bound, fb, lists));
  } else {
    RVP_TRY(batch_blocks<NB>(seg_normals_k<0, true>, n_bound, &qb, &strided));
    if (strided)
      RVP_LAUNCH("seg_normals_k", s, seg_normals_k<0, true><<<qb, QT, 0, s>>>(A, k + 1, nullptr, h0, NB,
                                                                              rec.as<float>(), n_bound, fb, lists));
    else
      RVP_LAUNCH("seg_normals_k", s, seg_normals_k<0, false><<<qb, QT, 0, s>>>(A, k + 1, nullptr, h0, NB,
                                                                               rec.as<float>(), n_bound, fb, lists));
  }
  return solve_normals(rec.as<float>(), n_bound, G.off + nclouds, A.xs, A.ys, A.zs, A.perm, vpx, vpy, vpz, nx, ny,
                       nz, fb, nullptr, nullptr, nullptr, s);
}

int rvp_batch_normals_knn(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                          size_t n_bound, int k, float vpx, float vpy, float vpz, float *nx, float *ny, float *nz,
                          rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && nx && ny && nz, RVP_E_ARG, "null pointer");
  RVP_CHECK(nclouds > 0, RVP_E_ARG, "nclouds must be positive");
  RVP_CHECK(k >= 1 && k + 1 <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 63]");
  return batch_normals_knn_impl(x, y, z, offsets, nclouds, n_bound, k, NbrLists{}, vpx, vpy, vpz, nx, ny, nz, S(st));
}

int rvp_batch_normals_knn_lists(const float *x, const float *y, const float *z, const uint32_t *offsets,
                                int nclouds, size_t n_bound, int k, const int32_t *list_idx, const float *list_d2,
                                const int32_t *list_counts, int list_cap, float vpx, float vpy, float vpz, float *nx,
                                float *ny, float *nz, rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && nx && ny && nz, RVP_E_ARG, "null pointer");
  RVP_CHECK(list_idx && list_counts, RVP_E_ARG, "null radius list");
  RVP_CHECK(nclouds > 0, RVP_E_ARG, "nclouds must be positive");
  RVP_CHECK(k >= 1 && k + 1 <= MAX_TOPK, RVP_E_UNSUPPORTED, "k must be in [1, 63]");
  RVP_CHECK(list_cap >= 1 && list_cap <= MAX_TOPK, RVP_E_ARG, "list_cap must be in [1, 64]");
  NbrLists lists;
  // rows shorter than k + 1 entries can serve no query: every query is searched
  if (k + 1 <= list_cap) {
    lists.idx = list_idx;
    lists.d2 = list_d2;
    lists.cnt = list_counts;
    lists.cap = list_cap;
  }
  return batch_normals_knn_impl(x, y, z, offsets, nclouds, n_bound, k, lists, vpx, vpy, vpz, nx, ny, nz, S(st));
}

int rvp_batch_normals_radius(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                             size_t n_bound, const float *radius, const float *cell, int max_nn, float vpx, float vpy,
                             float vpz, float *nx, float *ny, float *nz, rvp_stream_t st) {
  RVP_CHECK(x && y && z && offsets && radius && nx && ny && nz, RVP_E_ARG, "null pointer");
  RVP_CHECK(nclouds > 0, RVP_E_ARG, "nclouds must be positive");
  RVP_CHECK(max_nn >= 1 && max_nn <= MAX_TOPK, RVP_E_UNSUPPORTED, "max_nn must be in [1, 64]");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  SegGrid G;
  RVP_TRY(G.build(x, y, z, offsets, nclouds, n_bound, false, 0, cell ? cell : radius, s));
  unsigned long long *fb = nullptr;
  RVP_TRY(fallback_counter(s, &fb));
  Tmp rec;
  RVP_TRY(rec.alloc((size_t)REC_WORDS * n_bound * sizeof(float), s));
  // radius lists: batches of consecutive queries, whose shared windows stage once in LDS (faster than the
  // strided batches that balance k-NN queries, also for small counts)
  const int nb = radius_normals_batch();
  const unsigned qb = ceil_div(ceil_div(n_bound, (size_t)nb), QW);
  const SegArrays A = G.arrays();
  RVP_LAUNCH("seg_normals_k", s, seg_normals_k<1, false><<<qb, QT, 0, s>>>(A, max_nn, radius, 1, nb,
                                                                           rec.as<float>(), n_bound, fb,
                                                                           NbrLists{}));
  return solve_normals(rec.as<float>(), n_bound, G.off + nclouds, A.xs, A.ys, A.zs, A.perm, vpx, vpy, vpz, nx, ny,
                       nz, fb, nullptr, nullptr, nullptr, s);
}

int rvp_normals_fallback_count(size_t *count) {
  RVP_CHECK(count, RVP_E_ARG, "null argument");
  *count = 0;
  const LastNormalsCall last = g_last_normals;
  if (!last.counter) return RVP_OK;
  // read on the call's own stream and device; the caller's current device is restored
  int cur = 0;
  RVP_HIP(hipGetDevice(&cur));
  if (cur != last.dev) RVP_HIP(hipSetDevice(last.dev));
  unsigned long long h = 0;
  hipError_t e = hipMemcpyAsync(&h, last.counter, sizeof(h), hipMemcpyDeviceToHost, last.s);
  if (e == hipSuccess) e = hipStreamSynchronize(last.s);
  if (cur != last.dev) (void)hipSetDevice(cur);
  if (e != hipSuccess) return fail(RVP_E_DEVICE, std::string("reading the fallback counter: ") + hipGetErrorString(e));
  *count = (size_t)h;
  return RVP_OK;
}

}  // extern "C"
