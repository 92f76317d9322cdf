// prims.hip -- exclusive scan, stable LSD radix sort (8-bit digits, wave-ballot ranking), compaction.
#include "prims.hpp"

#include <utility>

namespace rvp {

// ------------------------------------------------------------------------------------------------
// exclusive scan (reduce -> scan partials -> downsweep)
// ------------------------------------------------------------------------------------------------
template <class Out>
__global__ __launch_bounds__(PRIM_THREADS) void scan_reduce_k(const uint32_t *__restrict__ in, size_t n,
                                                               Out *__restrict__ partials) {
  __shared__ Out sm[PRIM_THREADS / WAVE];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  Out acc = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    size_t i = base + (size_t)j * PRIM_THREADS + threadIdx.x;
    if (i < n) acc += (Out)in[i];
  }
  acc = wave_sum(acc);
  if (lane() == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    Out t = 0;
#pragma unroll
    for (int i = 0; i < PRIM_THREADS / WAVE; ++i) t += sm[i];
    partials[blockIdx.x] = t;
  }
}

template <class Out>
__global__ __launch_bounds__(PRIM_THREADS) void scan_partials_k(Out *__restrict__ partials, size_t nb,
                                                                 Out *__restrict__ total) {
  __shared__ Out sm[PRIM_THREADS / WAVE];
  Out carry = 0;
  for (size_t base = 0; base < nb; base += PRIM_THREADS) {
    size_t i = base + threadIdx.x;
    Out v = i < nb ? partials[i] : (Out)0;
    Out bt;
    Out ex = block_excl_scan_256(v, bt, sm);
    if (i < nb) partials[i] = carry + ex;
    carry += bt;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class Out>
__global__ __launch_bounds__(PRIM_THREADS) void scan_down_k(const uint32_t *in, Out *out, size_t n,
                                                             const Out *__restrict__ partials) {
  __shared__ Out sm[PRIM_THREADS / WAVE];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE;
  Out carry = partials[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    size_t i = base + (size_t)j * PRIM_THREADS + threadIdx.x;
    Out v = i < n ? (Out)in[i] : (Out)0;
    Out bt;
    Out ex = block_excl_scan_256(v, bt, sm);
    if (i < n) out[i] = carry + ex;
    carry += bt;
  }
}

template <class Out>
static int exclusive_scan_impl(const uint32_t *in, Out *out, size_t n, Out *total_dev, hipStream_t s) {
  if (n == 0) {
    if (total_dev) RVP_HIP(hipMemsetAsync(total_dev, 0, sizeof(Out), s));
    return RVP_OK;
  }
  const size_t nb = ceil_div(n, SCAN_TILE);
  Tmp part;
  RVP_TRY(part.alloc(nb * sizeof(Out), s));
  RVP_LAUNCH("scan_reduce_k", s, scan_reduce_k<Out><<<(unsigned)nb, PRIM_THREADS, 0, s>>>(in, n, part.as<Out>()));
  RVP_LAUNCH("scan_partials_k", s, scan_partials_k<Out><<<1, PRIM_THREADS, 0, s>>>(part.as<Out>(), nb, total_dev));
  RVP_LAUNCH("scan_down_k", s, scan_down_k<Out><<<(unsigned)nb, PRIM_THREADS, 0, s>>>(in, out, n, part.as<Out>()));
  return RVP_OK;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, hipStream_t s) {
  return exclusive_scan_impl<uint32_t>(in, out, n, total_dev, s);
}
int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *out, size_t n, uint64_t *total_dev, hipStream_t s) {
  return exclusive_scan_impl<uint64_t>(in, out, n, total_dev, s);
}

// ------------------------------------------------------------------------------------------------
// radix sort: per pass, a digit histogram per tile (digit-major), an exclusive scan over it, and a
// stable scatter that ranks equal digits inside each wave with 8 ballots (match-any).
// ------------------------------------------------------------------------------------------------
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = PRIM_THREADS * RS_ITEMS;

// element count of a device-sized sort: min(*n_dev, n) (n = the host bound the launch was sized for)
__device__ __forceinline__ size_t dev_count(size_t n, const uint32_t *n_dev) {
  return n_dev ? min((size_t)*n_dev, n) : n;
}
// a pass whose digit lies above the device-side key width only copies (the ping-pong stays in step)
__device__ __forceinline__ bool copy_pass(int shift, const int *bits_dev) { return bits_dev && shift >= *bits_dev; }

template <class K>
__global__ __launch_bounds__(PRIM_THREADS) void radix_hist_k(const K *__restrict__ keys, size_t n, int shift,
                                                              uint32_t *__restrict__ hist, unsigned nb,
                                                              const uint32_t *__restrict__ n_dev,
                                                              const int *__restrict__ bits_dev) {
  if (copy_pass(shift, bits_dev)) return;
  n = dev_count(n, n_dev);
  __shared__ uint32_t h[PRIM_THREADS / WAVE][256];
  const int t = threadIdx.x, w = t >> 6;
#pragma unroll
  for (int j = 0; j < PRIM_THREADS / WAVE; ++j) h[j][t] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    size_t i = base + (size_t)j * PRIM_THREADS + t;
    if (i < n) atomicAdd(&h[w][(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(size_t)t * nb + blockIdx.x] = h[0][t] + h[1][t] + h[2][t] + h[3][t];
}

template <class K>
__global__ __launch_bounds__(PRIM_THREADS) void radix_scatter_k(const K *__restrict__ kin,
                                                                 const uint32_t *__restrict__ vin,
                                                                 K *__restrict__ kout, uint32_t *__restrict__ vout,
                                                                 size_t n, int shift,
                                                                 const uint32_t *__restrict__ offs, unsigned nb,
                                                                 const uint32_t *__restrict__ n_dev,
                                                                 const int *__restrict__ bits_dev) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[PRIM_THREADS / WAVE][256];
  const int t = threadIdx.x, w = t >> 6;
  n = dev_count(n, n_dev);
  const size_t base = (size_t)blockIdx.x * RS_TILE;
  if (copy_pass(shift, bits_dev)) {
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
      const size_t i = base + (size_t)j * PRIM_THREADS + t;
      if (i < n) {
        kout[i] = kin[i];
        vout[i] = vin[i];
      }
    }
    return;
  }
  run[t] = offs[(size_t)t * nb + blockIdx.x];
  for (int j = 0; j < RS_ITEMS; ++j) {
    const size_t i = base + (size_t)j * PRIM_THREADS + t;
    const bool valid = i < n;
    K k = valid ? kin[i] : (K)0;
    uint32_t v = valid ? vin[i] : 0u;
    uint32_t d = (uint32_t)(k >> shift) & 255u;
#pragma unroll
    for (int q = 0; q < PRIM_THREADS / WAVE; ++q) wc[q][t] = 0;
    __syncthreads();
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const int rank = popc(peers & lanemask_lt());
    if (valid && rank == 0) wc[w][d] = (uint32_t)popc(peers);
    __syncthreads();
    {
      uint32_t s = run[t];
      uint32_t c0 = wc[0][t], c1 = wc[1][t], c2 = wc[2][t], c3 = wc[3][t];
      wc[0][t] = s; s += c0;
      wc[1][t] = s; s += c1;
      wc[2][t] = s; s += c2;
      wc[3][t] = s; s += c3;
      run[t] = s;
    }
    __syncthreads();
    if (valid) {
      const uint32_t pos = wc[w][d] + (uint32_t)rank;
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
  }
}

template <class K>
static int radix_sort_impl(K **keys, uint32_t **vals, K **kt, uint32_t **vt, size_t n, int bits, hipStream_t s,
                           const uint32_t *n_dev, const int *bits_dev) {
  if (n <= 1 || bits <= 0) return RVP_OK;
  RVP_CHECK(n < 0xffffffffull, RVP_E_UNSUPPORTED, "radix sort: more than 2^32-1 elements");
  const unsigned nb = ceil_div(n, RS_TILE);
  Tmp hist;
  RVP_TRY(hist.alloc((size_t)256 * nb * sizeof(uint32_t), s));
  for (int shift = 0; shift < bits; shift += 8) {
    RVP_LAUNCH("radix_hist_k", s, radix_hist_k<K><<<nb, PRIM_THREADS, 0, s>>>(*keys, n, shift, hist.as<uint32_t>(), nb,
                                                                             n_dev, bits_dev));
    RVP_TRY(exclusive_scan_u32(hist.as<uint32_t>(), hist.as<uint32_t>(), (size_t)256 * nb, nullptr, s));
    RVP_LAUNCH("radix_scatter_k", s, radix_scatter_k<K><<<nb, PRIM_THREADS, 0, s>>>(
                                         *keys, *vals, *kt, *vt, n, shift, hist.as<uint32_t>(), nb, n_dev, bits_dev));
    std::swap(*keys, *kt);
    std::swap(*vals, *vt);
  }
  return RVP_OK;
}

int radix_sort_pairs_u32(uint32_t **keys, uint32_t **vals, uint32_t **kt, uint32_t **vt, size_t n, int bits,
                         hipStream_t s, const uint32_t *n_dev, const int *bits_dev) {
  return radix_sort_impl<uint32_t>(keys, vals, kt, vt, n, bits, s, n_dev, bits_dev);
}
int radix_sort_pairs_u64(uint64_t **keys, uint32_t **vals, uint64_t **kt, uint32_t **vt, size_t n, int bits,
                         hipStream_t s, const uint32_t *n_dev, const int *bits_dev) {
  return radix_sort_impl<uint64_t>(keys, vals, kt, vt, n, bits, s, n_dev, bits_dev);
}

// ------------------------------------------------------------------------------------------------
// compaction
// ------------------------------------------------------------------------------------------------
__global__ void flag_eq_k(const uint8_t *__restrict__ flags, uint8_t want, size_t n, uint32_t *__restrict__ f) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = flags[i] == want ? 1u : 0u;
}

__global__ void scatter_aos_k(const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
                              const uint8_t *__restrict__ flags, uint8_t want, size_t n,
                              const uint32_t *__restrict__ pos, float *__restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i] == want) {
    size_t p = pos[i];
    out[3 * p + 0] = x[i];
    out[3 * p + 1] = y[i];
    out[3 * p + 2] = z[i];
  }
}

int compact_xyz_aos_pos(const float *x, const float *y, const float *z, const uint8_t *flags, uint8_t want, size_t n,
                        float *out_aos, uint32_t *pos, uint32_t *count_dev, hipStream_t s) {
  if (n == 0) {
    RVP_HIP(hipMemsetAsync(count_dev, 0, sizeof(uint32_t), s));
    return RVP_OK;
  }
  const unsigned nb = ceil_div(n, 256);
  RVP_LAUNCH("flag_eq_k", s, flag_eq_k<<<nb, 256, 0, s>>>(flags, want, n, pos));
  RVP_TRY(exclusive_scan_u32(pos, pos, n, count_dev, s));
  if (out_aos) RVP_LAUNCH("scatter_aos_k", s, scatter_aos_k<<<nb, 256, 0, s>>>(x, y, z, flags, want, n, pos, out_aos));
  return RVP_OK;
}

int compact_xyz_aos(const float *x, const float *y, const float *z, const uint8_t *flags, uint8_t want, size_t n,
                    float *out_aos, uint32_t *count_dev, hipStream_t s) {
  if (n == 0) {
    RVP_HIP(hipMemsetAsync(count_dev, 0, sizeof(uint32_t), s));
    return RVP_OK;
  }
  Tmp f;
  RVP_TRY(f.alloc(n * sizeof(uint32_t), s));
  const unsigned nb = ceil_div(n, 256);
  RVP_LAUNCH("flag_eq_k", s, flag_eq_k<<<nb, 256, 0, s>>>(flags, want, n, f.as<uint32_t>()));
  RVP_TRY(exclusive_scan_u32(f.as<uint32_t>(), f.as<uint32_t>(), n, count_dev, s));
  if (out_aos) {
    RVP_LAUNCH("scatter_aos_k", s, scatter_aos_k<<<nb, 256, 0, s>>>(x, y, z, flags, want, n, f.as<uint32_t>(), out_aos));
  }
  return RVP_OK;
}

}  // namespace rvp

