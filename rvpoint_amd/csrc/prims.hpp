// prims.hpp -- device-wide primitives: exclusive scan, stable LSD radix sort, stream compaction.
#pragma once

#include "common.hpp"

namespace rvp {

constexpr int PRIM_THREADS = 256;
constexpr int SCAN_ITEMS = 8;                       // items per thread per scan tile
constexpr int SCAN_TILE = PRIM_THREADS * SCAN_ITEMS; // 2048

// control words of a stream's look-back state (64-bit): the ticket word, then WORK_QUEUE_WORDS 32-bit counters
constexpr int CTL_WORDS = 192;
constexpr int WORK_QUEUE_WORDS = 2 * (CTL_WORDS - 1);
// The work counters of the persistent query kernels on stream s (zeroed when created; each launch leaves them
// zeroed), in stream order like the look-back state.
int stream_work_queue(hipStream_t s, uint32_t **queue);

// zeroes `words` (<= 256) 32-bit words with a kernel in stream order (capture-safe, unlike hipMemsetAsync)
int zero_words(void *p, uint32_t words, hipStream_t s);

// Exclusive scan out[i] = sum(in[0..i)), i < n; if total_dev != nullptr it receives sum(in[0..n)).
// In/Out may alias.  Out is uint32_t or uint64_t (int64_t via reinterpretation).
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, hipStream_t s);
// the same over the first min(*n_dev + add, n) elements (n_dev on the device, may be null)
int exclusive_scan_u32_bounded(const uint32_t *in, uint32_t *out, size_t n, const uint32_t *n_dev, uint32_t add,
                               uint32_t *total_dev, hipStream_t s);
// look-back waits of the single-pass scan that hit their spin limit since load (~0 if unreadable)
unsigned long long scan_spin_errors();
int exclusive_scan_u32_u64(const uint32_t *in, uint64_t *out, size_t n, uint64_t *total_dev, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `bits` bits of the key (8-bit digits).
// On return the sorted data is in (*keys, *vals): the function swaps the pointers with the temporaries
// when an odd number of passes leaves it there.  Device-sized form: n_dev (may be null) holds the element
// count (<= n, the bound the launches are sized for); bits_dev (may be null) the key width, passes above
// it only copy.  Neither is read on the host.  large_tiles: 4096-key tiles instead of 2048, for callers whose
// n is about the real count and at least half a million.
int radix_sort_pairs_u32(uint32_t **keys, uint32_t **vals, uint32_t **keys_tmp, uint32_t **vals_tmp,
                         size_t n, int bits, hipStream_t s, const uint32_t *n_dev = nullptr,
                         const int *bits_dev = nullptr, bool large_tiles = false);
int radix_sort_pairs_u64(uint64_t **keys, uint32_t **vals, uint64_t **keys_tmp, uint32_t **vals_tmp,
                         size_t n, int bits, hipStream_t s, const uint32_t *n_dev = nullptr,
                         const int *bits_dev = nullptr, bool large_tiles = false);

// Heads of the equal-key runs of sorted keys in one pass: vid[i] (i < n) = runs starting before i,
// start[r] = first index of run r, *total_dev = run count; n_dev (may be null) bounds the keys on the device.
int run_heads(const uint32_t *keys, size_t n, const uint32_t *n_dev, uint32_t *vid, uint32_t *start,
              uint32_t *total_dev, hipStream_t s);
int run_heads(const uint64_t *keys, size_t n, const uint32_t *n_dev, uint32_t *vid, uint32_t *start,
              uint32_t *total_dev, hipStream_t s);

// Stable compaction of points whose flag equals `want` into AoS (float[3*m]); *count_dev = m (device).
int compact_xyz_aos(const float *x, const float *y, const float *z, const uint8_t *flags, uint8_t want,
                    size_t n, float *out_aos, uint32_t *count_dev, hipStream_t s);

// X1 (ransac_plane.cpp:374-382) in one pass: points with |((a*x + b*y) + c*z) + d| <= t (as dist >= -t &&
// dist <= t) to in_aos, the rest to out_aos, each in index order (either may be null); counts_dev[0..1] =
// (inliers, outliers) on the device.  m = {a, b, c, d} on the host.
int plane_split_aos(const float *x, const float *y, const float *z, size_t n, const float *m, float t, float *in_aos,
                    float *out_aos, uint32_t *counts_dev, hipStream_t s);

// The reference's sequential fp32 sum s = fl(s + x_i) over i = 0..n-1 of non-negative terms, exact, in
// parallel (seqsum.hip).  mode 0: x_i = v[i]; mode 1: x_i = fl(fl(v[i] - *mean_dev)^2).  *out_dev = s.
int exact_seq_sum(const float *v, size_t n, int mode, const float *mean_dev, float *out_dev, hipStream_t s);
// The same for nseg independent segments [seg_off[s], seg_off[s+1]) (device offsets, n_bound >= seg_off[nseg]);
// mode 1 reads one mean per segment from mean_dev[s]; out[s] receives segment s's sum.  No host sync.
int exact_seq_sum_seg(const float *v, const uint32_t *seg_off, int nseg, size_t n_bound, int mode,
                      const float *mean_dev, float *out, hipStream_t s);
// Both sums of the SOR statistics per segment, one chunking: sum_out[s] = the sequential sum of v, and
// sq_out[s] = the sequential sum of fl(fl(v[i] - mean)^2) with mean = fl(sum_out[s] / len(s)).
int exact_seq_mean_var_seg(const float *v, const uint32_t *seg_off, int nseg, size_t n_bound, float *sum_out,
                           float *sq_out, hipStream_t s);

// Stable compaction in one pass of the points i < min(*n_dev, n) (n_dev may be null) whose u32 flag in pos is 1:
// pos is scanned in place into the output positions, out_aos (may be null) receives the flagged points,
// *count_dev = their count (device).
int compact_flagged(const float *x, const float *y, const float *z, uint32_t *pos, size_t n, const uint32_t *n_dev,
                    float *out_aos, uint32_t *count_dev, hipStream_t s);
// Stable compaction of points whose u8 flag equals 1, given those flags also as u32 in pos (scanned in place
// into the output positions); *count_dev = kept count (device).
int compact_xyz_aos_flagged(const float *x, const float *y, const float *z, const uint8_t *flags, size_t n,
                            float *out_aos, uint32_t *pos, uint32_t *count_dev, hipStream_t s);

// SOR's keep rule and stable compaction over a batch of ncl clouds [off[c], off[c+1]) in one launch: point i is
// kept iff means[i] <= fl(sums[c]/n_c) + fl(alpha * sqrtf(fl(sqs[c]/n_c))); keep (may be null) gets the u8 flags
// of all n_bound slots, out_aos the kept points in index order, out_off[0..ncl] each cloud's first output row and
// the kept total.  off must be a clamped partition (off[ncl] <= n_bound); nothing is read on the host.
int seg_keep_compact(const float *x, const float *y, const float *z, const float *means, const uint32_t *off,
                     int ncl, size_t n_bound, const float *sums, const float *sqs, float alpha, uint8_t *keep,
                     float *out_aos, uint32_t *out_off, hipStream_t s);

inline int bits_for(uint64_t max_value) {
  int b = 0;
  while (b < 64 && (max_value >> b) != 0) ++b;
  return b;
}

// block-wide exclusive scan for 256-thread blocks; smem must hold 4 T
template <class T> __device__ __forceinline__ T block_excl_scan_256(T v, T &block_total, T *smem) {
  T inc = wave_incl_scan(v);
  const int w = threadIdx.x >> 6;
  if (lane() == WAVE - 1) smem[w] = inc;
  __syncthreads();
  T wofs = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < PRIM_THREADS / WAVE; ++i) {
    T t = smem[i];
    if (i < w) wofs += t;
    tot += t;
  }
  __syncthreads();
  block_total = tot;
  return wofs + inc - v;
}

}  // namespace rvp
