// runtime.hip -- error state, memory pool set-up and the thin runtime helpers of the C ABI.
#include <mutex>

#include "common.hpp"
#include "prims.hpp"

namespace rvp {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// Keep freed pool memory cached (release threshold = max) so per-call temporaries do not go back to
// the driver at every synchronisation.
int pool_init() {
  static std::mutex mu;
  static bool done[64] = {false};
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return RVP_OK;
  std::lock_guard<std::mutex> lk(mu);
  if (done[dev]) return RVP_OK;
  hipMemPool_t pool;
  RVP_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
  uint64_t thr = ~0ull;
  RVP_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
  done[dev] = true;
  return RVP_OK;
}

}  // namespace rvp

using namespace rvp;

extern "C" {

const char *rvp_last_error(void) { return g_err.c_str(); }

const char *rvp_version(void) { return "rvpoint-amd 0.1 (gfx950)"; }

int rvp_device_count(int *count) {
  RVP_CHECK(count, RVP_E_ARG, "null argument");
  RVP_HIP(hipGetDeviceCount(count));
  return RVP_OK;
}

int rvp_set_device(int device) {
  RVP_HIP(hipSetDevice(device));
  return pool_init();
}

int rvp_malloc(void **ptr, size_t bytes) {
  RVP_CHECK(ptr, RVP_E_ARG, "null argument");
  RVP_TRY(pool_init());
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
  if (e != hipSuccess) return fail(RVP_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return RVP_OK;
}

int rvp_free(void *ptr) {
  if (ptr) RVP_HIP(hipFree(ptr));
  return RVP_OK;
}

int rvp_memcpy_h2d(void *dst, const void *src, size_t bytes, rvp_stream_t s) {
  if (!bytes) return RVP_OK;
  RVP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)));
  RVP_HIP(hipStreamSynchronize(S(s)));
  return RVP_OK;
}

int rvp_memcpy_d2h(void *dst, const void *src, size_t bytes, rvp_stream_t s) {
  if (!bytes) return RVP_OK;
  RVP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)));
  RVP_HIP(hipStreamSynchronize(S(s)));
  return RVP_OK;
}

int rvp_memcpy_d2d(void *dst, const void *src, size_t bytes, rvp_stream_t s) {
  if (!bytes) return RVP_OK;
  RVP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)));
  return RVP_OK;
}

int rvp_memset(void *dst, int value, size_t bytes, rvp_stream_t s) {
  if (!bytes) return RVP_OK;
  RVP_HIP(hipMemsetAsync(dst, value, bytes, S(s)));
  return RVP_OK;
}

int rvp_stream_create(rvp_stream_t *s) {
  RVP_CHECK(s, RVP_E_ARG, "null argument");
  RVP_TRY(pool_init());
  hipStream_t st;
  RVP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  *s = st;
  return RVP_OK;
}

int rvp_stream_destroy(rvp_stream_t s) {
  if (s) RVP_HIP(hipStreamDestroy(S(s)));
  return RVP_OK;
}

int rvp_stream_sync(rvp_stream_t s) {
  RVP_HIP(hipStreamSynchronize(S(s)));
  return RVP_OK;
}

int rvp_device_sync(void) {
  RVP_HIP(hipDeviceSynchronize());
  return RVP_OK;
}

int rvp_event_create(void **ev) {
  RVP_CHECK(ev, RVP_E_ARG, "null argument");
  hipEvent_t e;
  RVP_HIP(hipEventCreate(&e));
  *ev = e;
  return RVP_OK;
}

int rvp_event_destroy(void *ev) {
  if (ev) RVP_HIP(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return RVP_OK;
}

int rvp_event_record(void *ev, rvp_stream_t s) {
  RVP_HIP(hipEventRecord(static_cast<hipEvent_t>(ev), S(s)));
  return RVP_OK;
}

int rvp_event_elapsed(void *a, void *b, float *ms) {
  RVP_CHECK(ms, RVP_E_ARG, "null argument");
  RVP_HIP(hipEventSynchronize(static_cast<hipEvent_t>(b)));
  RVP_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
  return RVP_OK;
}

int rvp_exclusive_scan(const int32_t *counts, size_t n, int64_t *offsets, int64_t *total_host, rvp_stream_t s) {
  RVP_CHECK(offsets || n == 0, RVP_E_ARG, "null offsets");
  RVP_TRY(pool_init());
  RVP_TRY(exclusive_scan_u32_u64(reinterpret_cast<const uint32_t *>(counts), reinterpret_cast<uint64_t *>(offsets),
                                 n, reinterpret_cast<uint64_t *>(offsets) + n, S(s)));
  if (total_host) {
    RVP_HIP(hipMemcpyAsync(total_host, offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost, S(s)));
    RVP_HIP(hipStreamSynchronize(S(s)));
  }
  return RVP_OK;
}

}  // extern "C"
