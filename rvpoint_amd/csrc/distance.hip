// distance.hip -- D1 / D2 helpers of the drop-in API: get_dist_sq_rvv and get_inds_in_radius_rvv
// (src/core/rvv_common.cpp:9-107, scalar branches :36-43 and :55-68).
#include "prims.hpp"

namespace rvp {

__global__ void dist_sq_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                          size_t n, float qx, float qy, float qz, float *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = dist2(x[i], y[i], z[i], qx, qy, qz);
}

__global__ void subset_flags_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                               const int32_t *__restrict__ sub, size_t m, float qx, float qy, float qz, float r2,
                               float *__restrict__ d2, uint32_t *__restrict__ f) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int32_t j = sub[i];
  const float d = dist2(x[j], y[j], z[j], qx, qy, qz);
  d2[i] = d;
  f[i] = d <= r2 ? 1u : 0u;
}

__global__ void subset_scatter_k(const int32_t *__restrict__ sub, size_t m, const float *__restrict__ d2,
                                 const uint32_t *__restrict__ pos, float r2, int32_t *__restrict__ oi,
                                 float *__restrict__ od) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !(d2[i] <= r2)) return;
  const uint32_t p = pos[i];
  oi[p] = sub[i];
  od[p] = d2[i];
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_dist_sq(const float *x, const float *y, const float *z, size_t n, float qx, float qy, float qz, float *out,
                rvp_stream_t st) {
  if (n == 0) return RVP_OK;
  RVP_CHECK(x && y && z && out, RVP_E_ARG, "null pointer");
  RVP_LAUNCH("dist_sq_k", S(st), dist_sq_k<<<ceil_div(n, 256), 256, 0, S(st)>>>(x, y, z, n, qx, qy, qz, out));
  RVP_HIP(hipStreamSynchronize(S(st)));
  return RVP_OK;
}

int rvp_inds_in_radius(const float *x, const float *y, const float *z, const int32_t *subset, size_t m, float qx,
                       float qy, float qz, float r2, int32_t *out_idx, float *out_d2, size_t *n_out,
                       rvp_stream_t st) {
  RVP_CHECK(n_out, RVP_E_ARG, "null n_out");
  *n_out = 0;
  if (m == 0) return RVP_OK;
  RVP_CHECK(x && y && z && subset && out_idx && out_d2, RVP_E_ARG, "null pointer");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  Tmp d2, f, cnt;
  RVP_TRY(d2.alloc(m * sizeof(float), s));
  RVP_TRY(f.alloc(m * sizeof(uint32_t), s));
  RVP_TRY(cnt.alloc(sizeof(uint32_t), s));
  const unsigned nb = ceil_div(m, 256);
  RVP_LAUNCH("subset_flags_k", s, subset_flags_k<<<nb, 256, 0, s>>>(x, y, z, subset, m, qx, qy, qz, r2,
                                                                    d2.as<float>(), f.as<uint32_t>()));
  RVP_TRY(exclusive_scan_u32(f.as<uint32_t>(), f.as<uint32_t>(), m, cnt.as<uint32_t>(), s));
  RVP_LAUNCH("subset_scatter_k", s, subset_scatter_k<<<nb, 256, 0, s>>>(subset, m, d2.as<float>(), f.as<uint32_t>(),
                                                                        r2, out_idx, out_d2));
  uint32_t c = 0;
  RVP_HIP(hipMemcpyAsync(&c, cnt.p, sizeof(c), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *n_out = c;
  return RVP_OK;
}

}  // extern "C"
