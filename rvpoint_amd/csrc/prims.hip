// prims.hip -- exclusive scan, stable LSD radix sort (8-bit digits, wave-ballot ranking), compaction.
#include "prims.hpp"

#include <map>
#include <mutex>
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

// ------------------------------------------------------------------------------------------------
// single-pass exclusive scan (decoupled look-back) for 32-bit sums.  Tiles take their index from a ticket
// counter in dispatch order (atomicInc wraps it back to 0 after the last tile), publish their aggregate,
// and wave 0 looks back over up to 64 predecessors at a time for the nearest inclusive prefix.  A tile only
// waits on tiles with smaller tickets, which were dispatched before it, so the wait ends; a spin limit
// turns a wait that would not end into a counted error instead of a hang.  Status words carry the call's
// epoch (30 bits), so the per-stream status array never needs clearing.
// ------------------------------------------------------------------------------------------------
constexpr uint32_t SP_AGG = 1, SP_INC = 2;
constexpr uint32_t SP_SPIN_LIMIT = 1u << 24;
__device__ unsigned long long g_scan_spin_errors;

__device__ __forceinline__ uint64_t sp_word(uint32_t epoch, uint32_t flag, uint32_t v) {
  return ((uint64_t)((epoch << 2) | flag) << 32) | v;
}

// (in and out may be the same array: a tile reads all its inputs before it writes)
__global__ __launch_bounds__(PRIM_THREADS) void scan_1pass_k(const uint32_t *in, uint32_t *out, size_t n,
                                                              uint64_t *__restrict__ status,
                                                              uint32_t *__restrict__ ticket, uint32_t nb,
                                                              uint32_t epoch, uint32_t *__restrict__ total) {
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ uint32_t s_tile, s_prefix;
  const int t = threadIdx.x;
  if (t == 0) s_tile = atomicInc(ticket, nb - 1);
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t base = (size_t)tile * SCAN_TILE;
  uint32_t v[SCAN_ITEMS];
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = base + (size_t)j * PRIM_THREADS + t;
    v[j] = i < n ? in[i] : 0u;
    acc += v[j];
  }
  acc = wave_sum(acc);
  if (lane() == 0) sm[t >> 6] = acc;
  __syncthreads();
  const uint32_t agg = sm[0] + sm[1] + sm[2] + sm[3];
  if (t < WAVE) {
    const int l = t;
    uint32_t prefix = 0;
    if (tile == 0) {
      if (l == 0)
        __hip_atomic_store(&status[0], sp_word(epoch, SP_INC, agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (l == 0)
        __hip_atomic_store(&status[tile], sp_word(epoch, SP_AGG, agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      int64_t w0 = (int64_t)tile - 1;
      bool bad = false;
      for (;;) {
        const int64_t idx = w0 - l;
        uint64_t word = sp_word(epoch, SP_INC, 0);  // before tile 0: an inclusive zero
        uint32_t spins = 0;
        for (;;) {
          bool ready = true;
          if (idx >= 0) {
            word = __hip_atomic_load(&status[idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            ready = (uint32_t)(word >> 34) == epoch && ((word >> 32) & 3u) != 0u;
          }
          if (!ballot(!ready)) break;
          if (++spins >= SP_SPIN_LIMIT) {
            bad = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (bad) break;
        const uint64_t inc = ballot(((word >> 32) & 3u) == SP_INC);
        const uint32_t val = (uint32_t)word;
        if (inc) {
          const int first = __ffsll((unsigned long long)inc) - 1;  // nearest inclusive predecessor
          prefix += wave_sum(l <= first ? val : 0u);
          break;
        }
        prefix += wave_sum(val);
        w0 -= WAVE;
      }
      if (bad && l == 0) atomicAdd(&g_scan_spin_errors, 1ull);
      if (l == 0)
        __hip_atomic_store(&status[tile], sp_word(epoch, SP_INC, prefix + agg), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (l == 0) {
      s_prefix = prefix;
      if (total && tile == nb - 1) *total = prefix + agg;
    }
  }
  __syncthreads();
  uint32_t carry = s_prefix;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = base + (size_t)j * PRIM_THREADS + t;
    uint32_t bt;
    const uint32_t ex = block_excl_scan_256(v[j], bt, sm);
    if (i < n) out[i] = carry + ex;
    carry += bt;
  }
}

// per-(device, stream) look-back state: status words, ticket counter, epoch
struct ScanState {
  uint64_t *status = nullptr;
  uint32_t *ticket = nullptr;
  size_t cap = 0;
  uint32_t epoch = 0;
};
static std::mutex g_scan_mu;
static std::map<std::pair<int, hipStream_t>, ScanState> g_scan;

static int scan_state(hipStream_t s, size_t tiles, uint64_t **status, uint32_t **ticket, uint32_t *epoch) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scan_mu);
  ScanState &st = g_scan[{dev, s}];
  if (!st.ticket) {
    RVP_HIP(hipMalloc(&st.ticket, sizeof(uint32_t)));
    RVP_HIP(hipMemset(st.ticket, 0, sizeof(uint32_t)));
  }
  if (st.cap < tiles) {
    // the stream may still run a scan on the old array: wait before replacing it
    RVP_HIP(hipStreamSynchronize(s));
    if (st.status) RVP_HIP(hipFree(st.status));
    const size_t cap = std::max<size_t>(tiles, 4096);
    RVP_HIP(hipMalloc(&st.status, cap * sizeof(uint64_t)));
    RVP_HIP(hipMemset(st.status, 0, cap * sizeof(uint64_t)));
    st.cap = cap;
  }
  st.epoch = (st.epoch + 1) & 0x3fffffffu;
  if (st.epoch == 0) st.epoch = 1;  // 0 is the cleared state
  *status = st.status;
  *ticket = st.ticket;
  *epoch = st.epoch;
  return RVP_OK;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, hipStream_t s) {
  if (n == 0) {
    if (total_dev) RVP_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint32_t), s));
    return RVP_OK;
  }
  const size_t nb = ceil_div(n, SCAN_TILE);
  RVP_CHECK(nb < 0xffffffffull, RVP_E_UNSUPPORTED, "scan too large");
  uint64_t *status = nullptr;
  uint32_t *ticket = nullptr, epoch = 0;
  RVP_TRY(scan_state(s, nb, &status, &ticket, &epoch));
  RVP_LAUNCH("scan_1pass_k", s, scan_1pass_k<<<(unsigned)nb, PRIM_THREADS, 0, s>>>(in, out, n, status, ticket,
                                                                                  (uint32_t)nb, epoch, total_dev));
  return RVP_OK;
}
int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *out, size_t n, uint64_t *total_dev, hipStream_t s) {
  return exclusive_scan_impl<uint64_t>(in, out, n, total_dev, s);
}

unsigned long long scan_spin_errors() {
  unsigned long long h = 0;
  if (hipMemcpyFromSymbol(&h, HIP_SYMBOL(g_scan_spin_errors), sizeof(h)) != hipSuccess) return ~0ull;
  return h;
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

