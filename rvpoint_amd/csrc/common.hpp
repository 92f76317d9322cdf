// common.hpp -- shared host/device helpers for librvp_hip.so (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>

#include "../../include/rvp.h"

// Every fp32 operation of the reference path rounds once: no multiply-add contraction anywhere in the
// library (also passed as -ffp-contract=off).  Division and sqrtf are the correctly rounded forms
// (-fhip-fp32-correctly-rounded-divide-sqrt); HIP's __fsqrt_rn is an approximation and is not used.
#pragma clang fp contract(off)

namespace rvp {

// ------------------------------------------------------------------------------------------------
// error plumbing: every C entry point returns an int status; the message is kept per host thread.
// ------------------------------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define RVP_HIP(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return ::rvp::fail(RVP_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

#define RVP_CHECK(cond, code, msg)                                                                 \
  do {                                                                                             \
    if (!(cond)) return ::rvp::fail((code), (msg));                                               \
  } while (0)

#define RVP_TRY(expr)                                                                              \
  do {                                                                                             \
    int s_ = (expr);                                                                               \
    if (s_ != RVP_OK) return s_;                                                                   \
  } while (0)

// kernel launch check (catches bad launch configs; faults surface at the next sync)
#define RVP_LAUNCHED() RVP_HIP(hipGetLastError())

// ------------------------------------------------------------------------------------------------
// temporaries: a caching allocator keyed by (device, stream).  A block freed on a stream is handed out
// again only on that stream, so stream order makes the reuse safe without events; blocks stay cached until
// rvp_cache_trim (or an allocation failure) returns them to the driver.
// ------------------------------------------------------------------------------------------------
int pool_init();
void *cache_alloc(size_t bytes, hipStream_t s, size_t *cap);
void cache_free(void *p, size_t cap, hipStream_t s);

// A temporary comes from the cache and goes back to it on its stream; a persistent buffer (owned by a
// handle that may outlive the stream it was built on) uses hipMalloc/hipFree.  Buffers are reused when
// large enough.
struct Tmp {
  void *p = nullptr;
  size_t cap = 0;
  bool persistent = false;
  hipStream_t s = nullptr;
  Tmp() = default;
  Tmp(const Tmp &) = delete;
  Tmp &operator=(const Tmp &) = delete;
  ~Tmp() { release(); }
  void release() {
    if (p) {
      if (persistent) (void)hipFree(p);
      else cache_free(p, cap, s);
    }
    p = nullptr;
    cap = 0;
  }
  int alloc(size_t bytes, hipStream_t stream, bool keep = false) {
    if (bytes == 0) bytes = 16;
    if (p && keep == persistent && bytes <= cap && (keep || stream == s)) return RVP_OK;
    release();
    s = stream;
    persistent = keep;
    if (keep) {
      const hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) {
        p = nullptr;
        return fail(RVP_E_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
      }
      cap = bytes;
    } else {
      p = cache_alloc(bytes, stream, &cap);
      if (!p) return fail(RVP_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return RVP_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
  void swap(Tmp &o) {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
    std::swap(persistent, o.persistent);
    std::swap(s, o.s);
  }
};

inline hipStream_t S(rvp_stream_t s) { return static_cast<hipStream_t>(s); }

// ------------------------------------------------------------------------------------------------
// optional per-kernel timing: when enabled (rvp_profile_enable), every RVP_LAUNCH records a HIP event
// pair around the launch on its own stream; rvp_profile_collect sums the elapsed times per kernel.
// ------------------------------------------------------------------------------------------------
bool prof_enabled();
bool prof_want(const char *name);
// tuning knob read by the k-NN grid builders (rvp_tune_set "knn_cell_div")
double tune_knn_cell_div();
// dense-table cells per point a batched grid may use (rvp_tune_set "seg_cells_per_point")
double tune_seg_cells_per_point();
int tune_queries_per_wave();  // 0: chosen from the query count
int tune_work_steal();  // batched SOR runs: 0 strided, 1 per-XCD ranges with stealing, 2 per-XCD ranges strided
int tune_knn_start_cells();  // half-width in cells of an unbounded k-NN search's first box
int tune_normals_per_wave();  // consecutive queries per wave of the radius normals (0: the default)
// blocks of `threads` threads of `kernel` the current device holds at once (occupancy x CUs), cached
int resident_blocks(const void *kernel, int threads, unsigned *blocks);

// Pinned host staging for small device-to-host read-backs, one growing buffer per (device, stream), held
// (locked) by this object until it goes out of scope: copies into it are true async DMA, and callers that
// share a stream take turns.
struct PinnedStage {
  struct Slot;
  Slot *slot = nullptr;
  int get(hipStream_t s, size_t bytes, void **host);
  ~PinnedStage();
};
hipEvent_t prof_start(hipStream_t s);
void prof_stop(const char *name, hipStream_t s, hipEvent_t start);

#define RVP_LAUNCH(name, stream, ...)                                                              \
  do {                                                                                             \
    hipEvent_t ev_ = (::rvp::prof_enabled() && ::rvp::prof_want(name)) ? ::rvp::prof_start(stream) \
                                                                       : nullptr;                  \
    __VA_ARGS__;                                                                                   \
    if (ev_) ::rvp::prof_stop(name, stream, ev_);                                                  \
    RVP_LAUNCHED();                                                                                \
  } while (0)

constexpr int WAVE = 64;

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ------------------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane() { return (int)__lane_id(); }

// This wave's index within its block as a scalar: the compiler does not treat threadIdx.x / 64 as wave-uniform,
// and a per-lane copy turns every per-query value derived from it into vector registers and divergent branches
// (one query kernel: 79 VGPRs with the per-lane index, 53 with this).
__device__ __forceinline__ int wave_in_block() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

__device__ __forceinline__ int popc(uint64_t m) { return __popcll(m); }

__device__ __forceinline__ uint64_t lanemask_lt() { return (uint64_t)__lanemask_lt(); }

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// (d2, idx) -> 64-bit key whose unsigned order is the reference's std::pair<float,int> order
// (spatial_hashing.cpp:203-209); valid because d2 is a sum of squares, never -0 or negative.
__device__ __forceinline__ uint64_t pack_key(float d2, uint32_t idx) {
  return ((uint64_t)fbits(d2) << 32) | (uint64_t)idx;
}
__device__ __forceinline__ float key_d2(uint64_t k) { return bitsf((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return (uint32_t)(k & 0xffffffffu); }

// D1 (rvv_common.cpp:36-43): ((dx*dx + dy*dy) + dz*dz), each op rounded once, in this order.
__device__ __forceinline__ float dist2(float px, float py, float pz, float qx, float qy, float qz) {
  float dx = __fsub_rn(px, qx);
  float dy = __fsub_rn(py, qy);
  float dz = __fsub_rn(pz, qz);
  return __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
}

// Lane permutes through ds_bpermute with the byte address computed here (lane index * 4, taken mod 64
// lanes), without the bounds handling of HIP's __shfl family: every caller's partner lies in the wave.
__device__ __forceinline__ uint32_t bperm(uint32_t v, int addr) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute(addr, (int)v);
}
__device__ __forceinline__ int bperm(int v, int addr) { return __builtin_amdgcn_ds_bpermute(addr, v); }
__device__ __forceinline__ float bperm(float v, int addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}
__device__ __forceinline__ uint64_t bperm(uint64_t v, int addr) {
  const uint32_t lo = bperm((uint32_t)v, addr), hi = bperm((uint32_t)(v >> 32), addr);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double bperm(double v, int addr) {
  return __longlong_as_double((long long)bperm((uint64_t)__double_as_longlong(v), addr));
}
__device__ __forceinline__ long long bperm(long long v, int addr) { return (long long)bperm((uint64_t)v, addr); }
// byte address of lane l ^ m, of lane l - d, of lane l + d (the last two wrap around the wave)
__device__ __forceinline__ int xor_addr(int m) { return (lane() ^ m) << 2; }
__device__ __forceinline__ int up_addr(int d) { return ((lane() - d) << 2) & 0xfc; }
__device__ __forceinline__ int down_addr(int d) { return ((lane() + d) << 2) & 0xfc; }
template <class T> __device__ __forceinline__ T shfl_xor(T v, int m) { return bperm(v, xor_addr(m)); }

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) { return shfl_xor(v, m); }
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) { return bperm(v, src << 2); }
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ float readlanef(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// a value every lane holds alike, moved to a scalar register
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ uint32_t readlaneu(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

template <class T> __device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const T t = bperm(v, up_addr(d));
    if (l >= d) v += t;
  }
  return v;
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = WAVE / 2; d > 0; d >>= 1) v += shfl_xor(v, d);
  return v;
}

// next representable float toward +inf / -inf (finite inputs)
__device__ __forceinline__ float next_up(float f) {
  uint32_t u = fbits(f);
  if (f == 0.0f) return bitsf(1u);
  return bitsf((int32_t)u >= 0 ? u + 1 : u - 1);
}
__device__ __forceinline__ float next_down(float f) { return -next_up(-f); }

}  // namespace rvp
