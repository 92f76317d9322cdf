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

// zeroes `words` 32-bit words in stream order (a kernel rather than hipMemsetAsync, so that a stream capture
// records no memset node)
__global__ void zero_words_k(uint32_t *__restrict__ p, uint32_t words) {
  if (threadIdx.x < words) p[threadIdx.x] = 0u;
}
int zero_words(void *p, uint32_t words, hipStream_t s) {
  RVP_CHECK(words <= 256, RVP_E_ARG, "zero_words: at most 256 words");
  RVP_LAUNCH("zero_words_k", s, zero_words_k<<<1, 256, 0, s>>>(static_cast<uint32_t *>(p), words));
  return RVP_OK;
}

template <class Out>
static int exclusive_scan_impl(const uint32_t *in, Out *out, size_t n, Out *total_dev, hipStream_t s) {
  if (n == 0) {
    if (total_dev) RVP_TRY(zero_words(total_dev, sizeof(Out) / 4, s));
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
// single-pass exclusive scan (decoupled look-back) for 32-bit sums.  Tiles take their index from the stream's
// control word in dispatch order (take_tile), publish their aggregate, and wave 0 looks back over up to 64
// predecessors at a time for the nearest inclusive prefix.  A tile only waits on tiles with smaller tickets,
// which were dispatched before it, so the wait ends; a spin limit turns a wait that would not end into a
// counted error instead of a hang.  Status words carry the launch's epoch (30 bits, kept on the device), so
// the per-stream status array is cleared only when it is created.  The value travels in the same 64-bit word
// as its flag, so the status accesses are relaxed device-scope atomics: no other memory has to become visible
// with them, and an acquire / release pair would write back and invalidate the XCD's L2 on every status access.
// ------------------------------------------------------------------------------------------------
constexpr uint32_t SP_AGG = 1, SP_INC = 2;
constexpr uint32_t SP_SPIN_LIMIT = 1u << 24;
__device__ unsigned long long g_scan_spin_errors;

__device__ __forceinline__ uint64_t sp_word(uint32_t epoch, uint32_t flag, uint32_t v) {
  return ((uint64_t)((epoch << 2) | flag) << 32) | v;
}

// Wave 0 of a tile: publish the tile's aggregate, look back for the exclusive prefix of the tiles before it,
// publish the inclusive value; returns the prefix on every lane of the wave.
__device__ uint32_t tile_lookback(uint64_t *__restrict__ status, uint32_t tile, uint32_t epoch, uint32_t agg) {
  const int l = lane();
  uint32_t prefix = 0;
  if (tile == 0) {
    if (l == 0)
      __hip_atomic_store(&status[0], sp_word(epoch, SP_INC, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  if (l == 0)
    __hip_atomic_store(&status[tile], sp_word(epoch, SP_AGG, agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int64_t w0 = (int64_t)tile - 1;
  bool bad = false;
  for (;;) {
    const int64_t idx = w0 - l;
    uint64_t word = sp_word(epoch, SP_INC, 0);  // before tile 0: an inclusive zero
    uint32_t spins = 0;
    for (;;) {
      bool ready = true;
      if (idx >= 0) {
        word = __hip_atomic_load(&status[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    __hip_atomic_store(&status[tile], sp_word(epoch, SP_INC, prefix + agg), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  return prefix;
}

// A tile index and the launch's epoch from the stream's control word ctl = (generation << 32) | ticket, in one
// atomic: tiles take tickets in dispatch order, and the taker of the last ticket (nb - 1) moves the word to the
// next generation with the ticket back at zero, so every tile of a launch reads the same generation and the next
// launch on the stream the next one.  The epoch of the status words is that generation mod 2^30 - 1, plus one
// (0 is the cleared state).  Kept on the device, the epoch needs no host argument, so a captured graph replays
// with fresh epochs.  nb on a counter left off zero (counted, the tile does nothing).
__device__ __forceinline__ uint32_t take_tile(uint64_t *ctl, uint32_t nb, uint32_t *s_tile, uint32_t *s_epoch) {
  if (threadIdx.x == 0) {
    const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long *>(ctl), 1ull);
    const uint32_t t = (uint32_t)old;
    if (t == nb - 1) atomicAdd(reinterpret_cast<unsigned long long *>(ctl), (1ull << 32) - (unsigned long long)nb);
    if (t >= nb) atomicAdd(&g_scan_spin_errors, 1ull);
    *s_tile = t < nb ? t : nb;
    *s_epoch = (uint32_t)((old >> 32) % 0x3fffffffull) + 1u;
  }
  __syncthreads();
  return *s_tile;
}

// One tile of a single-pass scan whose thread t holds the tile's values t*SCAN_ITEMS .. t*SCAN_ITEMS + 7 in v:
// v becomes the exclusive scan over the whole launch (the tile's look-back prefix included); returns the tile's
// aggregate.  *s_prefix receives the prefix (shared memory of the block).  Every thread of the block calls it.
__device__ __forceinline__ uint32_t tile_scan_items(uint32_t (&v)[SCAN_ITEMS], uint64_t *__restrict__ status,
                                                    uint32_t tile, uint32_t epoch, uint32_t *sm, uint32_t *s_prefix) {
  uint32_t run = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const uint32_t x = v[j];
    v[j] = run;
    run += x;
  }
  uint32_t agg;
  const uint32_t texcl = block_excl_scan_256(run, agg, sm);
  if (threadIdx.x < WAVE) {
    const uint32_t prefix = tile_lookback(status, tile, epoch, agg);
    if (threadIdx.x == 0) *s_prefix = prefix;
  }
  __syncthreads();
  const uint32_t base = *s_prefix + texcl;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) v[j] += base;
  return agg;
}

// Single-pass exclusive scan.  Thread t of a tile owns its elements t*SCAN_ITEMS .. t*SCAN_ITEMS + 7
// (consecutive, 16-byte vector accesses when VEC): a local scan, ONE block-wide scan of the 256 thread totals,
// the look-back, and the writes.  (Strided items would need a block scan per item, each with its barriers.)
// in and out may be the same array: a tile reads all its inputs before it writes.
// n_dev (may be null) bounds the element count on the device: nn = min(*n_dev + add, n); only the tiles nn
// needs take tickets, the rest of a launch sized for n leaves at once
template <bool VEC>
__global__ __launch_bounds__(PRIM_THREADS) void scan_1pass_k(const uint32_t *in, uint32_t *out, size_t n_bound,
                                                              const uint32_t *__restrict__ n_dev, uint32_t add,
                                                              uint64_t *__restrict__ status,
                                                              uint64_t *__restrict__ ctl,
                                                              uint32_t *__restrict__ total) {
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ uint32_t s_tile, s_prefix, s_epoch;
  const size_t n = n_dev ? min((size_t)*n_dev + add, n_bound) : n_bound;
  const uint32_t nb = (uint32_t)((n + SCAN_TILE - 1) / SCAN_TILE);
  if (nb == 0) {
    if (total && blockIdx.x == 0 && threadIdx.x == 0) *total = 0u;
    return;
  }
  if (blockIdx.x >= nb) return;
  const int t = threadIdx.x;
  const uint32_t tile = take_tile(ctl, nb, &s_tile, &s_epoch);
  const uint32_t epoch = s_epoch;
  if (tile >= nb) return;
  const size_t i0 = (size_t)tile * SCAN_TILE + (size_t)t * SCAN_ITEMS;
  static_assert(SCAN_ITEMS == 8, "two 16-byte vectors per thread");
  uint32_t v[SCAN_ITEMS];
  if (VEC && i0 + SCAN_ITEMS <= n) {
    const uint4 a = *reinterpret_cast<const uint4 *>(in + i0), b = *reinterpret_cast<const uint4 *>(in + i0 + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) v[j] = i0 + j < n ? in[i0 + j] : 0u;
  }
  const uint32_t agg = tile_scan_items(v, status, tile, epoch, sm, &s_prefix);
  if (total && tile == nb - 1 && t == 0) *total = s_prefix + agg;
  if (VEC && i0 + SCAN_ITEMS <= n) {
    *reinterpret_cast<uint4 *>(out + i0) = make_uint4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<uint4 *>(out + i0 + 4) = make_uint4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
      if (i0 + j < n) out[i0 + j] = v[j];
  }
}

// X1 split in one pass (ransac_plane.cpp:374-382): dist = ((a*x + b*y) + c*z) + d, inlier iff
// dist >= -t && dist <= t; inliers and outliers each compacted to AoS in index order.  A tile's inlier
// rank comes from the look-back; its outlier rank is its index minus that.
__global__ __launch_bounds__(PRIM_THREADS) void plane_split_k(const float *__restrict__ x,
                                                               const float *__restrict__ y,
                                                               const float *__restrict__ z, size_t n, float a,
                                                               float b, float c, float d, float th,
                                                               uint64_t *__restrict__ status,
                                                               uint64_t *__restrict__ ctl, uint32_t nb,
                                                               float *__restrict__ in_aos,
                                                               float *__restrict__ out_aos,
                                                               uint32_t *__restrict__ counts) {
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ uint32_t s_tile, s_prefix, s_epoch;
  const int t = threadIdx.x;
  const uint32_t tile = take_tile(ctl, nb, &s_tile, &s_epoch);
  const uint32_t epoch = s_epoch;
  if (tile >= nb) return;
  const size_t i0 = (size_t)tile * SCAN_TILE + (size_t)t * SCAN_ITEMS;
  float px[SCAN_ITEMS], py[SCAN_ITEMS], pz[SCAN_ITEMS];
  uint32_t f[SCAN_ITEMS], r[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    f[j] = 0u;
    px[j] = py[j] = pz[j] = 0.0f;
    if (i < n) {
      px[j] = x[i];
      py[j] = y[i];
      pz[j] = z[i];
      const float dist =
          __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(a, px[j]), __fmul_rn(b, py[j])), __fmul_rn(c, pz[j])), d);
      f[j] = (dist >= -th && dist <= th) ? 1u : 0u;
    }
    r[j] = f[j];
  }
  const uint32_t agg = tile_scan_items(r, status, tile, epoch, sm, &s_prefix);
  if (tile == nb - 1 && t == 0) {
    counts[0] = s_prefix + agg;
    counts[1] = (uint32_t)n - (s_prefix + agg);
  }
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    if (i < n) {
      float *o = f[j] ? in_aos : out_aos;
      const size_t q = f[j] ? (size_t)r[j] : i - (size_t)r[j];
      if (o) {
        o[3 * q] = px[j];
        o[3 * q + 1] = py[j];
        o[3 * q + 2] = pz[j];
      }
    }
  }
}

// Stable compaction of flagged points in one pass: flags in pos[i] (0 or 1) for i < nn = min(*n_dev, n);
// pos[i] becomes the number of flagged points before i, a flagged point i goes to out_aos[pos[i]], and
// *count = the number flagged.  The tile count comes from nn on the device: only those tiles take tickets,
// the rest of a launch sized for the bound n leaves at once.
__global__ __launch_bounds__(PRIM_THREADS) void compact_flagged_k(const float *__restrict__ x,
                                                                   const float *__restrict__ y,
                                                                   const float *__restrict__ z,
                                                                   uint32_t *__restrict__ pos, size_t n,
                                                                   const uint32_t *__restrict__ n_dev,
                                                                   float *__restrict__ out_aos,
                                                                   uint64_t *__restrict__ status,
                                                                   uint64_t *__restrict__ ctl,
                                                                   uint32_t *__restrict__ count) {
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ uint32_t s_tile, s_prefix, s_epoch;
  const size_t nn = n_dev ? min((size_t)*n_dev, n) : n;
  const uint32_t nb = (uint32_t)((nn + SCAN_TILE - 1) / SCAN_TILE);
  if (nb == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0u;
    return;
  }
  if (blockIdx.x >= nb) return;
  const int t = threadIdx.x;
  const uint32_t tile = take_tile(ctl, nb, &s_tile, &s_epoch);
  const uint32_t epoch = s_epoch;
  if (tile >= nb) return;
  const size_t i0 = (size_t)tile * SCAN_TILE + (size_t)t * SCAN_ITEMS;
  uint32_t f[SCAN_ITEMS], r[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) r[j] = f[j] = i0 + j < nn ? pos[i0 + j] : 0u;
  const uint32_t agg = tile_scan_items(r, status, tile, epoch, sm, &s_prefix);
  if (tile == nb - 1 && t == 0) *count = s_prefix + agg;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    if (i < nn) {
      pos[i] = r[j];
      if (f[j] && out_aos) {
        out_aos[3 * (size_t)r[j]] = x[i];
        out_aos[3 * (size_t)r[j] + 1] = y[i];
        out_aos[3 * (size_t)r[j] + 2] = z[i];
      }
    }
  }
}

// Heads of the equal-key runs of sorted keys[0, nn) (nn = min(*n_dev, n)) in one pass: vid[i] = number of
// runs starting before i (i < n; nn.. get the total), start[r] = first index of run r, *total = run count.
template <class K>
__global__ __launch_bounds__(PRIM_THREADS) void run_heads_k(const K *__restrict__ keys, size_t n,
                                                             const uint32_t *__restrict__ n_dev,
                                                             uint32_t *__restrict__ vid, uint32_t *__restrict__ start,
                                                             uint64_t *__restrict__ status,
                                                             uint64_t *__restrict__ ctl, uint32_t nb,
                                                             uint32_t *__restrict__ total) {
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ uint32_t s_tile, s_prefix, s_epoch;
  const int t = threadIdx.x;
  const uint32_t tile = take_tile(ctl, nb, &s_tile, &s_epoch);
  const uint32_t epoch = s_epoch;
  if (tile >= nb) return;
  const size_t nn = n_dev ? min((size_t)*n_dev, n) : n;
  const size_t i0 = (size_t)tile * SCAN_TILE + (size_t)t * SCAN_ITEMS;
  uint32_t f[SCAN_ITEMS], ex[SCAN_ITEMS];
  K prev = (i0 > 0 && i0 - 1 < nn) ? keys[i0 - 1] : (K)0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    const K k = i < nn ? keys[i] : (K)0;
    ex[j] = f[j] = (i < nn && (i == 0 || k != prev)) ? 1u : 0u;
    prev = k;
  }
  const uint32_t agg = tile_scan_items(ex, status, tile, epoch, sm, &s_prefix);
  if (tile == nb - 1 && t == 0) *total = s_prefix + agg;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    if (i < n) vid[i] = ex[j];
    if (f[j]) start[ex[j]] = (uint32_t)i;
  }
}

// Per-(device, stream) look-back state: status words and the control words: word 0 is (generation, ticket; see
// take_tile), the rest the work counters of the persistent query kernels (stream_work_queue).
// Launches on one stream run in stream order whichever host thread enqueued them, and each accepts only status
// words of its own epoch, so several threads may drive one stream.  A status array that has been handed out is
// never freed while the process runs: a larger one replaces it and the old one is kept (retired), because a
// launch enqueued by another thread may still read it.
struct ScanState {
  uint64_t *status = nullptr;
  uint64_t *ctl = nullptr;
  size_t cap = 0;
  std::vector<uint64_t *> retired;
};
static std::mutex g_scan_mu;
static std::map<std::pair<int, hipStream_t>, ScanState> g_scan;

// Clears fresh look-back state in the order of the stream that will use it.  A plain hipMemset is enqueued on
// the null stream, which does not order against non-blocking streams: the first look-back launch on such a
// stream could then read the allocation's previous contents (a nonzero ticket skips a tile; a stale status
// word of a matching epoch hands out a wrong prefix), and the sort scatters or gathers built on those prefixes
// would leave their arrays.
__global__ void state_clear_k(uint64_t *__restrict__ status, size_t words, uint64_t *__restrict__ ctl) {
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ctl && i0 < CTL_WORDS) ctl[i0] = 0ull;
  for (size_t i = i0; i < words; i += (size_t)gridDim.x * blockDim.x) status[i] = 0ull;
}

static int scan_state(hipStream_t s, size_t tiles, uint64_t **status, uint64_t **ctl) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scan_mu);
  ScanState &st = g_scan[{dev, s}];
  uint64_t *clear_ctl = nullptr;
  if (!st.ctl) {
    RVP_HIP(hipMalloc(&st.ctl, CTL_WORDS * sizeof(uint64_t)));
    clear_ctl = st.ctl;
  }
  uint64_t *fresh = nullptr;
  size_t cap = st.cap;
  if (st.cap < tiles) {
    cap = std::max<size_t>(std::max<size_t>(tiles, 4096), 2 * st.cap);
    RVP_HIP(hipMalloc(&fresh, cap * sizeof(uint64_t)));
  }
  if (fresh || clear_ctl) {
    const size_t words = fresh ? cap : 0;
    const unsigned nb = (unsigned)std::min<size_t>(1024, std::max<size_t>(1, ceil_div(words, PRIM_THREADS)));
    RVP_LAUNCH("state_clear_k", s, state_clear_k<<<nb, PRIM_THREADS, 0, s>>>(fresh, words, clear_ctl));
  }
  if (fresh) {
    if (st.status) st.retired.push_back(st.status);
    st.status = fresh;
    st.cap = cap;
  }
  *status = st.status;
  *ctl = st.ctl;
  return RVP_OK;
}

int stream_work_queue(hipStream_t s, uint32_t **queue) {
  uint64_t *status = nullptr, *ctl = nullptr;
  RVP_TRY(scan_state(s, 0, &status, &ctl));
  *queue = reinterpret_cast<uint32_t *>(ctl + 1);
  return RVP_OK;
}

int exclusive_scan_u32_bounded(const uint32_t *in, uint32_t *out, size_t n, const uint32_t *n_dev, uint32_t add,
                               uint32_t *total_dev, hipStream_t s) {
  if (n == 0) {
    if (total_dev) RVP_TRY(zero_words(total_dev, 1, s));
    return RVP_OK;
  }
  const size_t nb = ceil_div(n, SCAN_TILE);
  RVP_CHECK(nb < 0xffffffffull && n < 0xffffffffull, RVP_E_UNSUPPORTED, "scan too large");
  uint64_t *status = nullptr;
  uint64_t *ctl = nullptr;
  RVP_TRY(scan_state(s, nb, &status, &ctl));
  if ((((uintptr_t)in | (uintptr_t)out) & 15u) == 0)
    RVP_LAUNCH("scan_1pass_k", s, scan_1pass_k<true><<<(unsigned)nb, PRIM_THREADS, 0, s>>>(in, out, n, n_dev, add,
                                                                                          status, ctl, total_dev));
  else
    RVP_LAUNCH("scan_1pass_k", s, scan_1pass_k<false><<<(unsigned)nb, PRIM_THREADS, 0, s>>>(in, out, n, n_dev, add,
                                                                                           status, ctl, total_dev));
  return RVP_OK;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, hipStream_t s) {
  return exclusive_scan_u32_bounded(in, out, n, nullptr, 0, total_dev, s);
}
int plane_split_aos(const float *x, const float *y, const float *z, size_t n, const float *m, float t, float *in_aos,
                    float *out_aos, uint32_t *counts_dev, hipStream_t s) {
  if (n == 0) {
    RVP_TRY(zero_words(counts_dev, 2, s));
    return RVP_OK;
  }
  const size_t nb = ceil_div(n, SCAN_TILE);
  RVP_CHECK(n < 0xffffffffull, RVP_E_UNSUPPORTED, "plane split: more than 2^32-1 points");
  uint64_t *status = nullptr;
  uint64_t *ctl = nullptr;
  RVP_TRY(scan_state(s, nb, &status, &ctl));
  RVP_LAUNCH("plane_split_k", s, plane_split_k<<<(unsigned)nb, PRIM_THREADS, 0, s>>>(
                                     x, y, z, n, m[0], m[1], m[2], m[3], t, status, ctl, (uint32_t)nb,
                                     in_aos, out_aos, counts_dev));
  return RVP_OK;
}

template <class K>
static int run_heads_impl(const K *keys, size_t n, const uint32_t *n_dev, uint32_t *vid, uint32_t *start,
                          uint32_t *total_dev, hipStream_t s) {
  if (n == 0) {
    RVP_TRY(zero_words(total_dev, 1, s));
    return RVP_OK;
  }
  RVP_CHECK(n < 0xffffffffull, RVP_E_UNSUPPORTED, "run heads: more than 2^32-1 keys");
  const size_t nb = ceil_div(n, SCAN_TILE);
  uint64_t *status = nullptr;
  uint64_t *ctl = nullptr;
  RVP_TRY(scan_state(s, nb, &status, &ctl));
  RVP_LAUNCH("run_heads_k", s, run_heads_k<K><<<(unsigned)nb, PRIM_THREADS, 0, s>>>(keys, n, n_dev, vid, start, status,
                                                                                  ctl, (uint32_t)nb,
                                                                                  total_dev));
  return RVP_OK;
}
int run_heads(const uint32_t *keys, size_t n, const uint32_t *n_dev, uint32_t *vid, uint32_t *start,
              uint32_t *total_dev, hipStream_t s) {
  return run_heads_impl(keys, n, n_dev, vid, start, total_dev, s);
}
int run_heads(const uint64_t *keys, size_t n, const uint32_t *n_dev, uint32_t *vid, uint32_t *start,
              uint32_t *total_dev, hipStream_t s) {
  return run_heads_impl(keys, n, n_dev, vid, start, total_dev, s);
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
// radix sort, one sweep per digit (after Adinets & Merrill, "Onesweep", 2022): one kernel histograms
// every digit position in a single read of the keys; each pass is then ONE kernel in which a tile ranks
// its keys (stable, waves own consecutive sub-ranges), publishes its per-digit counts, looks back over
// its predecessors' per-digit status words for its global offsets, stages the keys in LDS in digit order
// and writes them out in runs.  Status words are the scan's: (epoch, flag, value) in one relaxed 64-bit
// word, tiles ordered by a ticket so a tile waits only on tiles already running.
// ------------------------------------------------------------------------------------------------
constexpr int RS_ITEMS = 8;                       // keys per thread of a pass tile (16 for large sorts)
constexpr int RS_TILE = PRIM_THREADS * RS_ITEMS;  // 2048 keys per tile
constexpr int RS_MAX_PASSES = 8;
constexpr int RS_LOOKBACK = 8;  // predecessor status words one look-back step reads at once

// element count of a device-sized sort: min(*n_dev, n) (n = the host bound the launch was sized for)
__device__ __forceinline__ size_t dev_count(size_t n, const uint32_t *n_dev) {
  return n_dev ? min((size_t)*n_dev, n) : n;
}
// digit positions that carry key bits: all `passes`, or those below the device-side key width
__device__ __forceinline__ int active_passes(int passes, const int *bits_dev) {
  return bits_dev ? min(passes, (max(*bits_dev, 0) + 7) / 8) : passes;
}

__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-block digit histograms of every active position, stored (not added) to part[block][p][256]: global
// atomics on 256 * passes shared counters serialise across the blocks.
template <class K>
__global__ __launch_bounds__(PRIM_THREADS) void radix_hist_all_k(const K *__restrict__ keys, size_t n, int passes,
                                                                  uint32_t *__restrict__ part,
                                                                  uint32_t *__restrict__ ghist,
                                                                  const uint32_t *__restrict__ n_dev,
                                                                  const int *__restrict__ bits_dev) {
  __shared__ uint32_t h[RS_MAX_PASSES][256];
  const int t = threadIdx.x;
  const int act = active_passes(passes, bits_dev);
  n = dev_count(n, n_dev);
  for (int p = 0; p < act; ++p) h[p][t] = 0;
  if (blockIdx.x == 0)  // the reduction that follows adds into ghist
    for (int p = 0; p < act; ++p) ghist[p * 256 + t] = 0;
  __syncthreads();
  for (size_t i0 = (size_t)blockIdx.x * PRIM_THREADS; i0 < n; i0 += (size_t)gridDim.x * PRIM_THREADS) {
    const size_t i = i0 + t;
    const bool ok = i < n;
    const K k = ok ? keys[i] : (K)0;
    for (int p = 0; p < act; ++p) {
      const uint32_t d = (uint32_t)(k >> (8 * p)) & 255u;
      // a digit shared by the whole wave (the high digits of ordered input) is one add, not 64 conflicts
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
      const uint64_t okm = ballot(ok);
      if (!ballot(ok && d != d0)) {
        if (lane() == __ffsll((unsigned long long)okm) - 1) atomicAdd(&h[p][d0], (uint32_t)popc(okm));
      } else if (ok) {
        atomicAdd(&h[p][d], 1u);
      }
    }
  }
  __syncthreads();
  if ((size_t)blockIdx.x * PRIM_THREADS >= n) return;  // saw no key: the sum stops before this block
  for (int p = 0; p < act; ++p) part[((size_t)blockIdx.x * RS_MAX_PASSES + p) * 256 + t] = h[p][t];
}

// ghist[p][d] = sum over the hb partial histograms; block (p, q) sums blocks [q*RS_SUM, (q+1)*RS_SUM)
constexpr int RS_SUM = 16;
__global__ __launch_bounds__(PRIM_THREADS) void radix_hist_sum_k(const uint32_t *__restrict__ part, unsigned hb,
                                                                  size_t n, int passes, uint32_t *__restrict__ ghist,
                                                                  const uint32_t *__restrict__ n_dev,
                                                                  const int *__restrict__ bits_dev) {
  const int p = blockIdx.x % passes, q = blockIdx.x / passes, t = threadIdx.x;
  if (p >= active_passes(passes, bits_dev)) return;
  // histogram blocks that saw keys: the first ceil(n / PRIM_THREADS) of them
  const unsigned used = (unsigned)min((size_t)hb, (dev_count(n, n_dev) + PRIM_THREADS - 1) / PRIM_THREADS);
  const unsigned b0 = (unsigned)q * RS_SUM, b1 = min(used, b0 + RS_SUM);
  uint32_t v[RS_SUM];
#pragma unroll
  for (int j = 0; j < RS_SUM; ++j) v[j] = b0 + j < b1 ? part[((size_t)(b0 + j) * RS_MAX_PASSES + p) * 256 + t] : 0u;
  uint32_t acc = 0;
#pragma unroll
  for (int j = 0; j < RS_SUM; ++j) acc += v[j];
  if (acc) atomicAdd(&ghist[p * 256 + t], acc);
}

template <class K, int ITEMS>
__global__ __launch_bounds__(PRIM_THREADS) void radix_pass_k(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                              K *__restrict__ kout, uint32_t *__restrict__ vout,
                                                              size_t n, int pass, int passes,
                                                              const uint32_t *__restrict__ ghist,
                                                              uint64_t *__restrict__ status,
                                                              uint64_t *__restrict__ ctl, uint32_t nb,
                                                              const uint32_t *__restrict__ n_dev,
                                                              const int *__restrict__ bits_dev) {
  constexpr int TILE = PRIM_THREADS * ITEMS;
  __shared__ uint32_t s_tile, s_epoch;
  __shared__ uint32_t wcnt[PRIM_THREADS / WAVE][256];  // per-wave digit counters, then per-wave offsets
  __shared__ uint32_t lstart[256];                     // first slot of digit d in the staged tile
  __shared__ uint32_t gbase[256];                      // global position of staged slot 0 of digit d's run
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ K skey[TILE];
  __shared__ uint32_t sval[TILE];
  const int t = threadIdx.x, w = t >> 6, l = lane();
  n = dev_count(n, n_dev);
  const int act = active_passes(passes, bits_dev);
  if (pass >= act) {
    // no key bits left: the data stays put, except that one copy lands it in the buffer the remaining
    // (even or odd) number of pointer swaps ends on; uniform over the launch, so no tile takes a ticket
    if (pass == act && ((passes - act) & 1)) {
      for (size_t i = (size_t)blockIdx.x * PRIM_THREADS + t; i < n; i += (size_t)gridDim.x * PRIM_THREADS) {
        kout[i] = kin[i];
        vout[i] = vin[i];
      }
    }
    return;
  }
  // only the tiles the device-side count needs take tickets (the ticket returns to zero after exactly that
  // many); the launch's other blocks leave at once instead of queueing on the counter
  const uint32_t nb_act = (uint32_t)((n + TILE - 1) / TILE);
  if (blockIdx.x >= nb_act) return;
  const uint32_t gcount = ghist[pass * 256 + t];  // issued before the ticket round trip
#pragma unroll
  for (int q = 0; q < 4; ++q) wcnt[w][l + 64 * q] = 0;
  const uint32_t tile = take_tile(ctl, nb_act, &s_tile, &s_epoch);
  const uint32_t epoch = s_epoch;
  if (tile >= nb_act) return;
  const size_t base = (size_t)tile * TILE;
  const int shift = 8 * pass;
  // rank: wave w owns keys [base + w*512, base + (w+1)*512), in index order
  K k[ITEMS];
  uint32_t v[ITEMS], r[ITEMS];
  const size_t wb = base + (size_t)w * (ITEMS * WAVE);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const size_t i = wb + (size_t)j * WAVE + l;
    const bool ok = i < n;
    k[j] = ok ? kin[i] : (K)0;
    v[j] = ok ? vin[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const bool ok = wb + (size_t)j * WAVE + l < n;
    const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
    uint64_t peers = ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)popc(peers & lanemask_lt());
    const uint32_t pre = ok ? wcnt[w][d] : 0u;
    lds_wave_sync();
    if (ok && rank == 0) wcnt[w][d] = pre + (uint32_t)popc(peers);
    lds_wave_sync();
    r[j] = pre + rank;
  }
  __syncthreads();
  // thread t = digit t: per-wave offsets, the tile's count, its first staged slot
  uint32_t cnt;
  {
    const uint32_t c0 = wcnt[0][t], c1 = wcnt[1][t], c2 = wcnt[2][t], c3 = wcnt[3][t];
    cnt = c0 + c1 + c2 + c3;
    wcnt[0][t] = 0;
    wcnt[1][t] = c0;
    wcnt[2][t] = c0 + c1;
    wcnt[3][t] = c0 + c1 + c2;
  }
  if (tile == 0)
    __hip_atomic_store(&status[t], sp_word(epoch, SP_INC, cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    __hip_atomic_store(&status[(size_t)tile * 256 + t], sp_word(epoch, SP_AGG, cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  uint32_t tot;
  const uint32_t ls = block_excl_scan_256(cnt, tot, sm);
  const uint32_t gofs = block_excl_scan_256(gcount, tot, sm);
  // look back over the preceding tiles for the keys of digit t ahead of this tile, RS_LOOKBACK status words
  // in flight at a time: the walk consumes ready words in order up to the nearest inclusive one
  uint32_t excl = 0;
  if (tile > 0) {
    int64_t j = (int64_t)tile - 1;  // next predecessor to consume
    uint32_t spins = 0;
    bool done = false, bad = false;
    while (!done) {
      uint64_t wd[RS_LOOKBACK];
#pragma unroll
      for (int q = 0; q < RS_LOOKBACK; ++q)
        wd[q] = j - q >= 0 ? __hip_atomic_load(&status[(size_t)(j - q) * 256 + t], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : sp_word(epoch, SP_INC, 0);  // before tile 0: an inclusive zero
      int used = 0;
      bool stop = false;
#pragma unroll
      for (int q = 0; q < RS_LOOKBACK; ++q) {
        if (stop) continue;
        const uint32_t flag = (uint32_t)(wd[q] >> 32) & 3u;
        if ((uint32_t)(wd[q] >> 34) != epoch || flag == 0u) {
          stop = true;
          continue;
        }
        excl += (uint32_t)wd[q];
        ++used;
        if (flag == SP_INC) stop = done = true;
      }
      j -= used;
      if (!done && used == 0) {
        if (++spins >= SP_SPIN_LIMIT) {
          bad = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (bad) atomicAdd(&g_scan_spin_errors, 1ull);
    __hip_atomic_store(&status[(size_t)tile * 256 + t], sp_word(epoch, SP_INC, excl + cnt), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  lstart[t] = ls;
  gbase[t] = gofs + excl - ls;  // mod 2^32: slot s of digit t's run lands at gbase[t] + s
  __syncthreads();
  // stage in digit order, then write runs
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    if (wb + (size_t)j * WAVE + l < n) {
      const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
      const uint32_t s = lstart[d] + wcnt[w][d] + r[j];
      skey[s] = k[j];
      sval[s] = v[j];
    }
  }
  __syncthreads();
  const uint32_t m = (uint32_t)min((size_t)TILE, n - base);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const uint32_t s = (uint32_t)(j * PRIM_THREADS + t);
    if (s < m) {
      const K kk = skey[s];
      const uint32_t pos = gbase[(uint32_t)(kk >> shift) & 255u] + s;
      kout[pos] = kk;
      vout[pos] = sval[s];
    }
  }
}

template <class K>
static int radix_sort_impl(K **keys, uint32_t **vals, K **kt, uint32_t **vt, size_t n, int bits, hipStream_t s,
                           const uint32_t *n_dev, const int *bits_dev, bool large_tiles) {
  if (n <= 1 || bits <= 0) return RVP_OK;
  RVP_CHECK(n < 0xffffffffull, RVP_E_UNSUPPORTED, "radix sort: more than 2^32-1 elements");
  const int passes = std::min<int>((bits + 7) / 8, (int)sizeof(K));
  const unsigned hb = std::min(ceil_div(n, RS_TILE) * 2, 1024u);
  // large tiles: half the tiles, so half the look-back chain and its status traffic (worth it from about
  // half a million keys; below that the longer tile is the critical path)
  const unsigned nb = ceil_div(n, (size_t)PRIM_THREADS * (large_tiles ? 2 * RS_ITEMS : RS_ITEMS));
  Tmp gh, part;
  RVP_TRY(gh.alloc((size_t)passes * 256 * sizeof(uint32_t), s));
  RVP_TRY(part.alloc((size_t)hb * RS_MAX_PASSES * 256 * sizeof(uint32_t), s));
  RVP_LAUNCH("radix_hist_all_k", s, radix_hist_all_k<K><<<hb, PRIM_THREADS, 0, s>>>(*keys, n, passes,
                                                                                    part.as<uint32_t>(),
                                                                                    gh.as<uint32_t>(), n_dev,
                                                                                    bits_dev));
  RVP_LAUNCH("radix_hist_sum_k", s, radix_hist_sum_k<<<(unsigned)passes * ceil_div(hb, RS_SUM), PRIM_THREADS, 0, s>>>(
                                        part.as<uint32_t>(), hb, n, passes, gh.as<uint32_t>(), n_dev, bits_dev));
  for (int p = 0; p < passes; ++p) {
    uint64_t *status = nullptr;
    uint64_t *ctl = nullptr;
    RVP_TRY(scan_state(s, (size_t)nb * 256, &status, &ctl));
    if (large_tiles)
      RVP_LAUNCH("radix_pass_k", s, radix_pass_k<K, 2 * RS_ITEMS><<<nb, PRIM_THREADS, 0, s>>>(
                                        *keys, *vals, *kt, *vt, n, p, passes, gh.as<uint32_t>(), status, ctl, nb,
                                        n_dev, bits_dev));
    else
      RVP_LAUNCH("radix_pass_k", s, radix_pass_k<K, RS_ITEMS><<<nb, PRIM_THREADS, 0, s>>>(
                                        *keys, *vals, *kt, *vt, n, p, passes, gh.as<uint32_t>(), status, ctl, nb,
                                        n_dev, bits_dev));
    std::swap(*keys, *kt);
    std::swap(*vals, *vt);
  }
  return RVP_OK;
}

int radix_sort_pairs_u32(uint32_t **keys, uint32_t **vals, uint32_t **kt, uint32_t **vt, size_t n, int bits,
                         hipStream_t s, const uint32_t *n_dev, const int *bits_dev, bool large_tiles) {
  return radix_sort_impl<uint32_t>(keys, vals, kt, vt, n, bits, s, n_dev, bits_dev, large_tiles);
}
int radix_sort_pairs_u64(uint64_t **keys, uint32_t **vals, uint64_t **kt, uint32_t **vt, size_t n, int bits,
                         hipStream_t s, const uint32_t *n_dev, const int *bits_dev, bool large_tiles) {
  return radix_sort_impl<uint64_t>(keys, vals, kt, vt, n, bits, s, n_dev, bits_dev, large_tiles);
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

// SOR's keep rule and the stable compaction of a batch in one pass (statistical_outlier_removal.cpp:53-62): point
// i of cloud c is kept iff means[i] <= fl(sums[c] / n_c) + fl(alpha * sqrtf(fl(sqs[c] / n_c))).  keep[i] (may be
// null) receives the flag for every i < n_bound (0 past the batch), kept points go to out_aos in index order,
// out_off[c] = the first output row of cloud c and out_off[ncl] = the kept total.  The launch is sized for
// n_bound; only the tiles of off[ncl] points take tickets, the other blocks clear their share of keep.
__global__ __launch_bounds__(PRIM_THREADS) void seg_keep_compact_k(
    const float *__restrict__ x, const float *__restrict__ y, const float *__restrict__ z,
    const float *__restrict__ means, const uint32_t *__restrict__ off, int ncl, size_t n_bound,
    const float *__restrict__ sums, const float *__restrict__ sqs, float alpha, uint8_t *__restrict__ keep,
    float *__restrict__ out_aos, uint32_t *__restrict__ out_off, uint64_t *__restrict__ status,
    uint64_t *__restrict__ ctl) {
  __shared__ uint32_t sm[PRIM_THREADS / WAVE];
  __shared__ uint32_t s_tile, s_prefix, s_epoch;
  const int t = threadIdx.x;
  const uint32_t nn = off[ncl];  // <= n_bound (the offsets are a clamped partition)
  const uint32_t nb = (nn + SCAN_TILE - 1) / SCAN_TILE;
  if (blockIdx.x >= nb) {
    if (keep)
      for (size_t i = (size_t)blockIdx.x * SCAN_TILE + t; i < min((size_t)(blockIdx.x + 1) * SCAN_TILE, n_bound);
           i += PRIM_THREADS)
        keep[i] = 0;
    if (nb == 0 && blockIdx.x == 0)
      for (int c = t; c <= ncl; c += PRIM_THREADS) out_off[c] = 0u;
    return;
  }
  const uint32_t tile = take_tile(ctl, nb, &s_tile, &s_epoch);
  const uint32_t epoch = s_epoch;
  if (tile >= nb) return;
  const size_t i0 = (size_t)tile * SCAN_TILE + (size_t)t * SCAN_ITEMS;
  // the cloud of the thread's first point (largest c with off[c] <= i), then advanced point by point
  int c = 0;
  if (i0 < nn) {
    int lo = 0, hi = ncl - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= i0) lo = mid;
      else hi = mid - 1;
    }
    c = lo;
  }
  int cthr = -1;
  float thr = 0.0f;
  uint32_t f[SCAN_ITEMS], r[SCAN_ITEMS];
  int cl[SCAN_ITEMS];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    f[j] = 0u;
    cl[j] = -1;
    if (i < nn) {
      while (c + 1 < ncl && off[c + 1] <= i) ++c;
      if (c != cthr) {
        const float len = __uint2float_rn(off[c + 1] - off[c]);
        thr = __fadd_rn(__fdiv_rn(sums[c], len), __fmul_rn(alpha, sqrtf(__fdiv_rn(sqs[c], len))));
        cthr = c;
      }
      f[j] = means[i] <= thr ? 1u : 0u;
      cl[j] = c;
    }
    if (keep && i < n_bound) keep[i] = (uint8_t)f[j];
    r[j] = f[j];
  }
  const uint32_t agg = tile_scan_items(r, status, tile, epoch, sm, &s_prefix);
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const size_t i = i0 + j;
    if (f[j] && out_aos) {
      out_aos[3 * (size_t)r[j]] = x[i];
      out_aos[3 * (size_t)r[j] + 1] = y[i];
      out_aos[3 * (size_t)r[j] + 2] = z[i];
    }
    // a cloud's first point gives the cloud's first output row, also to the empty clouds just before it
    if (cl[j] >= 0 && off[cl[j]] == i)
      for (int e = cl[j]; e >= 0 && off[e] == i; --e) out_off[e] = r[j];
  }
  if (tile == nb - 1 && t == 0) {
    const uint32_t total = s_prefix + agg;
    for (int e = ncl; e >= 0 && (e == ncl || off[e] >= nn); --e) out_off[e] = total;
  }
}

int seg_keep_compact(const float *x, const float *y, const float *z, const float *means, const uint32_t *off,
                     int ncl, size_t n_bound, const float *sums, const float *sqs, float alpha, uint8_t *keep,
                     float *out_aos, uint32_t *out_off, hipStream_t s) {
  RVP_CHECK(n_bound < 0xffffffffull, RVP_E_UNSUPPORTED, "compaction: more than 2^32-1 points");
  const size_t nb = std::max<size_t>(1, ceil_div(n_bound, SCAN_TILE));
  uint64_t *status = nullptr;
  uint64_t *ctl = nullptr;
  RVP_TRY(scan_state(s, nb, &status, &ctl));
  RVP_LAUNCH("seg_keep_compact_k", s, seg_keep_compact_k<<<(unsigned)nb, PRIM_THREADS, 0, s>>>(
                                          x, y, z, means, off, ncl, n_bound, sums, sqs, alpha, keep, out_aos,
                                          out_off, status, ctl));
  return RVP_OK;
}

int compact_flagged(const float *x, const float *y, const float *z, uint32_t *pos, size_t n, const uint32_t *n_dev,
                    float *out_aos, uint32_t *count_dev, hipStream_t s) {
  if (n == 0) {
    RVP_TRY(zero_words(count_dev, 1, s));
    return RVP_OK;
  }
  RVP_CHECK(n < 0xffffffffull, RVP_E_UNSUPPORTED, "compaction: more than 2^32-1 points");
  const size_t nb = ceil_div(n, SCAN_TILE);
  uint64_t *status = nullptr;
  uint64_t *ctl = nullptr;
  RVP_TRY(scan_state(s, nb, &status, &ctl));
  RVP_LAUNCH("compact_flagged_k", s, compact_flagged_k<<<(unsigned)nb, PRIM_THREADS, 0, s>>>(
                                         x, y, z, pos, n, n_dev, out_aos, status, ctl, count_dev));
  return RVP_OK;
}

int compact_xyz_aos_flagged(const float *x, const float *y, const float *z, const uint8_t *flags, size_t n,
                            float *out_aos, uint32_t *pos, uint32_t *count_dev, hipStream_t s) {
  if (n == 0) {
    RVP_TRY(zero_words(count_dev, 1, s));
    return RVP_OK;
  }
  RVP_TRY(exclusive_scan_u32(pos, pos, n, count_dev, s));
  if (out_aos)
    RVP_LAUNCH("scatter_aos_k", s, scatter_aos_k<<<ceil_div(n, 256), 256, 0, s>>>(x, y, z, flags, 1, n, pos, out_aos));
  return RVP_OK;
}

int compact_xyz_aos(const float *x, const float *y, const float *z, const uint8_t *flags, uint8_t want, size_t n,
                    float *out_aos, uint32_t *count_dev, hipStream_t s) {
  if (n == 0) {
    RVP_TRY(zero_words(count_dev, 1, s));
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


// ------------------------------------------------------------------------------------------------
// Allocation-free plane split: the X1 split of rvp_plane_extract with its look-back state (status words,
// control word) and its two counts in a fixed per-device array instead of stream-ordered allocations, for
// callers that cannot allocate device memory (code under graph capture, serving loops).
// ------------------------------------------------------------------------------------------------
namespace rvp {

constexpr size_t NOALLOC_TILES = size_t(1) << 16;                  // status words: one per tile
constexpr size_t NOALLOC_MAX_POINTS = NOALLOC_TILES * SCAN_TILE;   // 134,217,728 points
// [0, NOALLOC_TILES): status words; NOALLOC_TILES: the control word (generation, ticket); NOALLOC_TILES + 1: the
// two counts
static __device__ __align__(256) uint64_t scratch[NOALLOC_TILES + 2];

static std::mutex g_noalloc_mu;
static uint64_t *g_noalloc[64];  // each device's copy of `scratch` (hipGetSymbolAddress)

// The control word's ticket must be zero when the split starts; a complete launch leaves it there, a failed one
// may not.  Moving to the next generation with a zero ticket also gives the call an epoch its predecessors did
// not use, so their status words are never taken for this call's.
__global__ void noalloc_ctl_reset_k(uint64_t *sc) {
  if (threadIdx.x == 0) sc[NOALLOC_TILES] = ((sc[NOALLOC_TILES] >> 32) + 1ull) << 32;
}

// this device's scratch address, looked up once per device with the device current
static int noalloc_scratch(uint64_t **ptr) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  RVP_CHECK(dev >= 0 && dev < 64, RVP_E_UNSUPPORTED, "device index");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    RVP_CHECK(e == hipSuccess && p, RVP_E_DEVICE,
              std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_noalloc[dev] = static_cast<uint64_t *>(p);
  }
  *ptr = g_noalloc[dev];
  return RVP_OK;
}

}  // namespace rvp

extern "C" int rvp_plane_extract_noalloc(const float *x, const float *y, const float *z, size_t n,
                                         const float *model_host, float dist_thresh, float *inliers_aos,
                                         size_t *n_in, float *outliers_aos, size_t *n_out, rvp_stream_t st) {
  using namespace rvp;
  RVP_CHECK(model_host && n_in && n_out, RVP_E_ARG, "null argument");
  *n_in = *n_out = 0;
  if (n == 0) return RVP_OK;
  RVP_CHECK(x && y && z, RVP_E_ARG, "null coordinate pointer");
  RVP_CHECK(n <= NOALLOC_MAX_POINTS, RVP_E_UNSUPPORTED,
            "rvp_plane_extract_noalloc: more points than its fixed scratch holds (134,217,728)");
  hipStream_t s = S(st);
  uint64_t *sc = nullptr;
  RVP_TRY(noalloc_scratch(&sc));
  const size_t nb = ceil_div(n, SCAN_TILE);
  uint32_t *counts = reinterpret_cast<uint32_t *>(sc + NOALLOC_TILES + 1);
  noalloc_ctl_reset_k<<<1, WAVE, 0, s>>>(sc);
  RVP_HIP(hipGetLastError());
  plane_split_k<<<(unsigned)nb, PRIM_THREADS, 0, s>>>(x, y, z, n, model_host[0], model_host[1], model_host[2],
                                                        model_host[3], dist_thresh, sc, sc + NOALLOC_TILES,
                                                        (uint32_t)nb, inliers_aos, outliers_aos, counts);
  RVP_HIP(hipGetLastError());
  uint32_t h[2] = {0, 0};
  RVP_HIP(hipMemcpyAsync(h, counts, sizeof(h), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *n_in = h[0];
  *n_out = h[1];
  return RVP_OK;
}
