// runtime.hip -- error state, memory pool set-up and the thin runtime helpers of the C ABI.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

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
int pool_init() { return RVP_OK; }

// ---- caching allocator ------------------------------------------------------------------------------
// size classes: 4 per octave above 4 KiB (at most 19% slack), 4 KiB below
static size_t size_class(size_t bytes) {
  if (bytes <= 4096) return 4096;
  int e = 63 - __builtin_clzll(bytes - 1);  // 2^e < bytes <= 2^(e+1)
  const size_t step = (size_t)1 << (e - 2);
  return (bytes + step - 1) / step * step;
}

struct CacheKey {
  int dev;
  hipStream_t s;
  size_t cls;
  bool operator<(const CacheKey &o) const {
    if (dev != o.dev) return dev < o.dev;
    if (s != o.s) return s < o.s;
    return cls < o.cls;
  }
};
static std::mutex g_cache_mu;
static std::map<CacheKey, std::vector<void *>> g_cache;

static void cache_release_all(int dev, bool only_stream, hipStream_t s) {
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto it = g_cache.begin(); it != g_cache.end();) {
      if (it->first.dev == dev && (!only_stream || it->first.s == s)) {
        drop.insert(drop.end(), it->second.begin(), it->second.end());
        it = g_cache.erase(it);
      } else {
        ++it;
      }
    }
  }
  if (drop.empty()) return;
  (void)hipDeviceSynchronize();
  for (void *p : drop) (void)hipFree(p);
}

// ---- capture arenas ---------------------------------------------------------------------------------
// While a stream is being captured into a graph, its temporaries come from an arena that the graph keeps: the
// graph's kernels hold their addresses for every replay, so those blocks must not go back to the shared cache
// (where eager work could take them) until the graph is destroyed.  Inside the capture a freed block returns to
// the arena and may serve a later stage of the same capture (the captured stream is one chain, so stages never
// overlap).  An arena takes blocks from the stream's cache first, so a capture preceded by an eager warm-up
// step allocates nothing.
struct CaptureArena {
  int dev = 0;
  hipStream_t s = nullptr;
  std::map<size_t, std::vector<void *>> free;   // by size class
  std::vector<std::pair<void *, size_t>> owned;  // every block, with its class
};
static std::map<std::pair<int, hipStream_t>, CaptureArena *> g_arenas;  // active captures (under g_cache_mu)

void *cache_alloc(size_t bytes, hipStream_t s, size_t *cap) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const size_t cls = size_class(bytes);
  CaptureArena *arena = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto ai = g_arenas.find({dev, s});
    if (ai != g_arenas.end()) {
      arena = ai->second;
      auto &fl = arena->free[cls];
      if (!fl.empty()) {
        void *p = fl.back();
        fl.pop_back();
        *cap = cls;
        return p;
      }
    }
    auto it = g_cache.find(CacheKey{dev, s, cls});
    if (it != g_cache.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      *cap = cls;
      if (arena) arena->owned.emplace_back(p, cls);
      return p;
    }
  }
  if (arena) {
    void *p = nullptr;
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    arena->owned.emplace_back(p, cls);
    *cap = cls;
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, cls) != hipSuccess) {
    (void)hipGetLastError();
    cache_release_all(dev, false, nullptr);  // give cached blocks back and retry once
    if (hipMalloc(&p, cls) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  *cap = cls;
  return p;
}

void cache_free(void *p, size_t cap, hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  auto ai = g_arenas.find({dev, s});
  if (ai != g_arenas.end()) {
    CaptureArena *a = ai->second;
    bool mine = false;
    for (auto &b : a->owned) mine = mine || b.first == p;
    if (!mine) a->owned.emplace_back(p, cap);  // allocated before the capture began: the graph keeps it too
    a->free[cap].push_back(p);
    return;
  }
  g_cache[CacheKey{dev, s, cap}].push_back(p);
}

static thread_local bool g_capturing = false;  // this thread is capturing a stream (no profiling events)

int arena_begin(hipStream_t s) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_cache_mu);
  RVP_CHECK(!g_arenas.count({dev, s}), RVP_E_STATE, "this stream is already being captured");
  CaptureArena *a = new CaptureArena();
  a->dev = dev;
  a->s = s;
  g_arenas[{dev, s}] = a;
  g_capturing = true;
  return RVP_OK;
}

CaptureArena *arena_end(hipStream_t s) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_capturing = false;
  auto ai = g_arenas.find({dev, s});
  if (ai == g_arenas.end()) return nullptr;
  CaptureArena *a = ai->second;
  g_arenas.erase(ai);
  return a;
}

// the graph is gone (its stream drained): the arena's blocks go back to that stream's cache
void arena_release(CaptureArena *a) {
  if (!a) return;
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (auto &b : a->owned) g_cache[CacheKey{a->dev, a->s, b.second}].push_back(b.first);
  }
  delete a;
}

// ---- kernel timer ---------------------------------------------------------------------------------
struct ProfRec {
  const char *name;
  hipEvent_t a, b;
};
static std::mutex g_prof_mu;
static std::vector<ProfRec> g_prof;
static std::vector<hipEvent_t> g_ev_pool;  // events reused across collects
static std::atomic<bool> g_prof_on{false};
static std::string g_prof_filter;          // empty: every kernel

bool prof_enabled() { return !g_capturing && g_prof_on.load(std::memory_order_relaxed); }

bool prof_want(const char *name) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  return g_prof_filter.empty() || g_prof_filter == name;
}

static hipEvent_t pool_event() {
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (!g_ev_pool.empty()) {
      hipEvent_t e = g_ev_pool.back();
      g_ev_pool.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static void release_event(hipEvent_t e) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_ev_pool.push_back(e);
}

hipEvent_t prof_start(hipStream_t s) {
  hipEvent_t e = pool_event();
  if (!e) return nullptr;
  if (hipEventRecord(e, s) != hipSuccess) {
    release_event(e);
    return nullptr;
  }
  return e;
}

void prof_stop(const char *name, hipStream_t s, hipEvent_t a) {
  hipEvent_t b = pool_event();
  if (!b || hipEventRecord(b, s) != hipSuccess) {
    release_event(a);
    if (b) release_event(b);
    return;
  }
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back(ProfRec{name, a, b});
}

static std::atomic<double> g_knn_cell_div{2.4};
static std::atomic<double> g_seg_cells_per_point{32.0};
static std::atomic<int> g_queries_per_wave{0};
static std::atomic<int> g_work_steal{0};
static std::atomic<int> g_knn_start_cells{1};
static std::atomic<int> g_normals_per_wave{0};
int tune_work_steal() { return g_work_steal.load(); }
int tune_knn_start_cells() { return g_knn_start_cells.load(); }
int tune_normals_per_wave() { return g_normals_per_wave.load(); }
double tune_knn_cell_div() { return g_knn_cell_div.load(); }
double tune_seg_cells_per_point() { return g_seg_cells_per_point.load(); }
int tune_queries_per_wave() { return g_queries_per_wave.load(); }

struct PinnedStage::Slot {
  std::mutex mu;
  void *p = nullptr;
  size_t cap = 0;
};
static std::mutex g_pin_mu;
static std::map<std::pair<int, hipStream_t>, PinnedStage::Slot *> g_pin;  // kept until the process ends

int PinnedStage::get(hipStream_t s, size_t bytes, void **host) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    Slot *&sl = g_pin[{dev, s}];
    if (!sl) sl = new Slot;
    slot = sl;
  }
  slot->mu.lock();
  if (slot->cap < bytes) {
    if (slot->p) RVP_HIP(hipHostFree(slot->p));
    slot->p = nullptr;
    slot->cap = 0;
    const size_t cap = std::max<size_t>(bytes, 64 << 10);
    RVP_HIP(hipHostMalloc(&slot->p, cap, hipHostMallocDefault));
    slot->cap = cap;
  }
  *host = slot->p;
  return RVP_OK;
}

PinnedStage::~PinnedStage() {
  if (slot) slot->mu.unlock();
}

static std::mutex g_occ_mu;
static std::map<std::tuple<int, const void *, int>, unsigned> g_occ;

int resident_blocks(const void *kernel, int threads, unsigned *blocks) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_occ_mu);
  const auto key = std::make_tuple(dev, kernel, threads);
  auto it = g_occ.find(key);
  if (it == g_occ.end()) {
    int per_cu = 0, cus = 0;
    RVP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    RVP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    it = g_occ.emplace(key, (unsigned)std::max(1, per_cu) * (unsigned)std::max(1, cus)).first;
  }
  *blocks = it->second;
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
  if (s) {
    int dev = 0;
    RVP_HIP(hipGetDevice(&dev));
    rvp::cache_release_all(dev, true, S(s));
    RVP_HIP(hipStreamDestroy(S(s)));
  }
  return RVP_OK;
}

int rvp_cache_trim(void) {
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  rvp::cache_release_all(dev, false, nullptr);
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

int rvp_stream_wait_event(rvp_stream_t s, void *ev) {
  RVP_CHECK(ev, RVP_E_ARG, "null event");
  RVP_HIP(hipStreamWaitEvent(S(s), static_cast<hipEvent_t>(ev), 0));
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

int rvp_exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, rvp_stream_t s) {
  RVP_CHECK((in && out) || n == 0, RVP_E_ARG, "null array");
  RVP_TRY(pool_init());
  return exclusive_scan_u32(in, out, n, total_dev, S(s));
}

int rvp_sort_pairs_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp, size_t n, int bits,
                       const uint32_t *n_dev, const int *bits_dev, int *in_tmp, rvp_stream_t s) {
  RVP_CHECK(in_tmp && ((keys && vals && keys_tmp && vals_tmp) || n == 0), RVP_E_ARG, "null argument");
  RVP_CHECK(bits >= 0 && bits <= 32, RVP_E_ARG, "bits must be in [0, 32]");
  RVP_TRY(pool_init());
  uint32_t *k = keys, *v = vals, *kt = keys_tmp, *vt = vals_tmp;
  RVP_TRY(radix_sort_pairs_u32(&k, &v, &kt, &vt, n, bits, S(s), n_dev, bits_dev, !n_dev && n >= (1u << 19)));
  *in_tmp = k != keys;
  return RVP_OK;
}

int rvp_sort_pairs_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_tmp, uint32_t *vals_tmp, size_t n, int bits,
                       const uint32_t *n_dev, const int *bits_dev, int *in_tmp, rvp_stream_t s) {
  RVP_CHECK(in_tmp && ((keys && vals && keys_tmp && vals_tmp) || n == 0), RVP_E_ARG, "null argument");
  RVP_CHECK(bits >= 0 && bits <= 64, RVP_E_ARG, "bits must be in [0, 64]");
  RVP_TRY(pool_init());
  uint64_t *k = keys, *kt = keys_tmp;
  uint32_t *v = vals, *vt = vals_tmp;
  RVP_TRY(radix_sort_pairs_u64(&k, &v, &kt, &vt, n, bits, S(s), n_dev, bits_dev, !n_dev && n >= (1u << 19)));
  *in_tmp = k != keys;
  return RVP_OK;
}

int rvp_scan_wait_errors(unsigned long long *count) {
  RVP_CHECK(count, RVP_E_ARG, "null argument");
  RVP_HIP(hipDeviceSynchronize());
  *count = scan_spin_errors();
  RVP_CHECK(*count != ~0ull, RVP_E_DEVICE, "reading the scan error counter failed");
  return RVP_OK;
}

int rvp_tune_set(const char *name, double value) {
  RVP_CHECK(name, RVP_E_ARG, "null name");
  if (std::string(name) == "knn_cell_div") {
    RVP_CHECK(value >= 0.25 && value <= 16.0, RVP_E_ARG, "knn_cell_div must be in [0.25, 16]");
    rvp::g_knn_cell_div.store(value);
    return RVP_OK;
  }
  if (std::string(name) == "seg_cells_per_point") {
    RVP_CHECK(value >= 1.0 && value <= 64.0, RVP_E_ARG, "seg_cells_per_point must be in [1, 64]");
    rvp::g_seg_cells_per_point.store(value);
    return RVP_OK;
  }
  if (std::string(name) == "queries_per_wave") {
    RVP_CHECK(value >= 0.0 && value <= 64.0 && value == (double)(int)value, RVP_E_ARG,
              "queries_per_wave must be an integer in [0, 64]");
    rvp::g_queries_per_wave.store((int)value);
    return RVP_OK;
  }
  if (std::string(name) == "work_steal") {
    RVP_CHECK(value == 0.0 || value == 1.0 || value == 2.0, RVP_E_ARG, "work_steal must be 0, 1 or 2");
    rvp::g_work_steal.store((int)value);
    return RVP_OK;
  }
  if (std::string(name) == "knn_start_cells") {
    RVP_CHECK(value >= 1.0 && value <= 8.0 && value == (double)(int)value, RVP_E_ARG,
              "knn_start_cells must be an integer in [1, 8]");
    rvp::g_knn_start_cells.store((int)value);
    return RVP_OK;
  }
  if (std::string(name) == "normals_per_wave") {
    RVP_CHECK(value >= 0.0 && value <= 64.0 && value == (double)(int)value, RVP_E_ARG,
              "normals_per_wave must be an integer in [0, 64]");
    rvp::g_normals_per_wave.store((int)value);
    return RVP_OK;
  }
  return fail(RVP_E_ARG, std::string("unknown tuning knob: ") + name);
}

int rvp_tune_get(const char *name, double *value) {
  RVP_CHECK(name && value, RVP_E_ARG, "null argument");
  if (std::string(name) == "knn_cell_div") {
    *value = rvp::g_knn_cell_div.load();
    return RVP_OK;
  }
  if (std::string(name) == "seg_cells_per_point") {
    *value = rvp::g_seg_cells_per_point.load();
    return RVP_OK;
  }
  if (std::string(name) == "queries_per_wave") {
    *value = (double)rvp::g_queries_per_wave.load();
    return RVP_OK;
  }
  if (std::string(name) == "work_steal") {
    *value = (double)rvp::g_work_steal.load();
    return RVP_OK;
  }
  if (std::string(name) == "knn_start_cells") {
    *value = (double)rvp::g_knn_start_cells.load();
    return RVP_OK;
  }
  if (std::string(name) == "normals_per_wave") {
    *value = (double)rvp::g_normals_per_wave.load();
    return RVP_OK;
  }
  return fail(RVP_E_ARG, std::string("unknown tuning knob: ") + name);
}

int rvp_profile_enable(int on) {
  g_prof_on.store(on != 0);
  return RVP_OK;
}

int rvp_profile_filter(const char *kernel) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof_filter = kernel ? kernel : "";
  return RVP_OK;
}

// text of records collected but not yet handed out (a size query leaves it here for the next call)
static std::string g_prof_text;

int rvp_profile_collect(char *buf, size_t buflen, size_t *needed) {
  std::vector<ProfRec> recs;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    recs.swap(g_prof);
  }
  std::map<std::string, std::pair<long, double>> acc;
  int status = RVP_OK;
  for (auto &r : recs) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto &e = acc[r.name];
      e.first += 1;
      e.second += ms;
    } else {
      status = fail(RVP_E_DEVICE, "profile: event query failed");
    }
    release_event(r.a);
    release_event(r.b);
  }
  std::string &out = g_prof_text;
  for (auto &kv : acc) {
    char line[256];
    snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
    out += line;
  }
  if (needed) *needed = out.size() + 1;
  if (buf && buflen > out.size()) {
    memcpy(buf, out.data(), out.size());
    buf[out.size()] = 0;
    out.clear();
  }
  return status;
}

// AoS (PointXYZ[n]) -> SoA, the copy the reference's drivers make between stages (pipeline_export.cpp:336-342)
}  // extern "C"

namespace rvp {
__global__ void aos_to_soa_k(const float *__restrict__ aos, size_t n, float *__restrict__ x, float *__restrict__ y,
                             float *__restrict__ z) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = aos[3 * i];
  y[i] = aos[3 * i + 1];
  z[i] = aos[3 * i + 2];
}
// grid-stride over min(*n_dev, n_bound) points
__global__ void aos_to_soa_bounded_k(const float *__restrict__ aos, size_t n_bound, const uint32_t *__restrict__ n_dev,
                                     float *__restrict__ x, float *__restrict__ y, float *__restrict__ z) {
  const size_t n = min((size_t)*n_dev, n_bound);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    x[i] = aos[3 * i];
    y[i] = aos[3 * i + 1];
    z[i] = aos[3 * i + 2];
  }
}
}  // namespace rvp

extern "C" int rvp_aos_to_soa(const float *aos, size_t n, float *x, float *y, float *z, rvp_stream_t st) {
  if (n == 0) return RVP_OK;
  RVP_CHECK(aos && x && y && z, RVP_E_ARG, "null pointer");
  RVP_LAUNCH("aos_to_soa", S(st), aos_to_soa_k<<<ceil_div(n, 256), 256, 0, S(st)>>>(aos, n, x, y, z));
  return RVP_OK;
}

extern "C" int rvp_aos_to_soa_bounded(const float *aos, size_t n_bound, const uint32_t *n_dev, float *x, float *y,
                                      float *z, rvp_stream_t st) {
  if (n_bound == 0) return RVP_OK;
  RVP_CHECK(aos && n_dev && x && y && z, RVP_E_ARG, "null pointer");
  // a quarter million threads: one pass for a frame's kept points, strided beyond
  const unsigned nb = (unsigned)std::min<size_t>(1024, ceil_div(n_bound, 256));
  RVP_LAUNCH("aos_to_soa_bounded", S(st), aos_to_soa_bounded_k<<<nb, 256, 0, S(st)>>>(aos, n_bound, n_dev, x, y, z));
  return RVP_OK;
}

// ---- graphs ------------------------------------------------------------------------------------------
struct rvp_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rvp::CaptureArena *arena = nullptr;
  hipStream_t s = nullptr;
};

extern "C" int rvp_graph_capture_begin(rvp_stream_t st) {
  using namespace rvp;
  RVP_CHECK(st, RVP_E_ARG, "graph capture needs a created stream (the null stream cannot be captured)");
  hipStream_t s = S(st);
  RVP_TRY(arena_begin(s));
  const hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    arena_release(arena_end(s));
    return fail(RVP_E_DEVICE, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  }
  return RVP_OK;
}

extern "C" int rvp_graph_capture_end(rvp_stream_t st, rvp_graph **out) {
  using namespace rvp;
  RVP_CHECK(st && out, RVP_E_ARG, "null argument");
  *out = nullptr;
  hipStream_t s = S(st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  CaptureArena *a = arena_end(s);
  if (e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    arena_release(a);
    return fail(RVP_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  const hipError_t e2 = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e2 != hipSuccess) {
    (void)hipGraphDestroy(g);
    arena_release(a);
    return fail(RVP_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
  }
  rvp_graph *r = new rvp_graph();
  r->graph = g;
  r->exec = x;
  r->arena = a;
  r->s = s;
  *out = r;
  return RVP_OK;
}

extern "C" int rvp_graph_launch(rvp_graph *g, rvp_stream_t st) {
  RVP_CHECK(g && g->exec, RVP_E_ARG, "null graph");
  RVP_CHECK(rvp::S(st) == g->s, RVP_E_ARG, "a graph replays on the stream it was captured on");
  RVP_HIP(hipGraphLaunch(g->exec, rvp::S(st)));
  return RVP_OK;
}

extern "C" int rvp_graph_destroy(rvp_graph *g) {
  if (!g) return RVP_OK;
  // the graph's last replay may still run: its blocks go back to the cache only after the stream drains
  const hipError_t e = hipStreamSynchronize(g->s);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  rvp::arena_release(g->arena);
  delete g;
  RVP_HIP(e);
  return RVP_OK;
}
