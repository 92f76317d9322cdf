// server.hip -- streaming frames through the batched pipeline with continuous batching.
//
// A server owns `slots` execution slots.  Each slot is a worker thread with its own HIP stream and its own
// device buffers for up to `slot_points` points.  Submitted frames wait in one FIFO queue; a slot that is
// free takes as many queued frames as fit its capacity (in arrival order, uneven sizes side by side), runs
// them as one batch -- voxel -> SOR -> radius normals -> RANSAC, one launch per stage for the whole batch,
// counts kept on the device between stages -- and hands every frame its own results.  Frames that arrive
// while other batches run are admitted by the next free slot, so batches of different slots overlap on
// the device.  Admission by HBM budget: the slots' working sets (an upper bound per point) must fit the
// budget, which fixes how many slots run at once.  Every frame's results equal what the batch of that
// frame alone returns, since each stage treats the clouds of a batch independently.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace rvp {

// device bytes one batch point may need at the peak of the pipeline (buffers the slot owns plus the
// temporaries of the batched stages: sort keys, grids with their dense cell tables, compaction)
constexpr size_t SERVER_BYTES_PER_POINT = 512;
constexpr int MAX_TOPK_API = 64;  // neighbour lists the batched stages keep per point

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Pinned host buffers, reused across frames: a submitted frame is copied once into one (the upload then runs as
// a true asynchronous copy), and results are copied from the device straight into others that the caller
// releases with rvp_result_free.  Process-wide (results may outlive their server); buffers go in power-of-two
// classes, at most POOL_KEEP bytes stay cached, the rest are freed on release.
class PinnedPool {
 public:
  float *get(size_t floats) {
    const size_t cls = class_of(floats * sizeof(float));
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto &fl = free_[cls];
      if (!fl.empty()) {
        float *p = fl.back();
        fl.pop_back();
        cached_ -= cls;
        return p;
      }
    }
    void *p = nullptr;
    if (hipHostMalloc(&p, cls, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu_);
    size_[p] = cls;
    return static_cast<float *>(p);
  }
  void put(float *p) {
    if (!p) return;
    bool drop = false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = size_.find(p);
      if (it == size_.end()) return;  // not a pool buffer
      if (cached_ + it->second <= POOL_KEEP) {
        free_[it->second].push_back(p);
        cached_ += it->second;
      } else {
        size_.erase(it);
        drop = true;
      }
    }
    if (drop) (void)hipHostFree(p);
  }

 private:
  static constexpr size_t POOL_KEEP = size_t(2) << 30;
  static size_t class_of(size_t bytes) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
  }
  std::mutex mu_;
  std::unordered_map<size_t, std::vector<float *>> free_;
  std::unordered_map<void *, size_t> size_;
  size_t cached_ = 0;
};
static PinnedPool g_pinned;

struct Request {
  uint64_t ticket;
  float *xyz = nullptr;  // AoS copy of the frame, pinned
  size_t n = 0;
  rvp_frame_result res{};
  bool done = false;
  int status = RVP_OK;
  std::string error;
  ~Request() { g_pinned.put(xyz); }
};

struct Slot {
  int dev = 0;
  hipStream_t s = nullptr;
  size_t cap = 0;
  // device: input AoS and SoA, voxel output and SoA, SOR output and SoA, normals (SoA, then AoS), per-cloud
  // parameters
  Tmp aos, x, y, z, vox, vx, vy, vz, sor, sx, sy, sz, nx, ny, nz, naos, off, voff, soff, status, leaf, radius;
  // pinned host staging of the per-frame offsets and parameters
  uint32_t *h_off = nullptr;
  std::thread th;
};

}  // namespace rvp

struct rvp_server {
  rvp_frame_params p{};
  std::vector<rvp::Slot *> slots;
  size_t slot_points = 0;
  int max_frames = 256;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<rvp::Request *> queue;
  std::unordered_map<uint64_t, rvp::Request *> live;
  uint64_t next_ticket = 1;
  bool stopping = false;
};

namespace rvp {

static int slot_alloc(Slot &S, int dev, size_t cap, int max_frames) {
  S.dev = dev;
  S.cap = cap;
  RVP_HIP(hipSetDevice(dev));
  RVP_HIP(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
  const size_t fb = cap * sizeof(float);
  for (Tmp *t : {&S.x, &S.y, &S.z, &S.vx, &S.vy, &S.vz, &S.sx, &S.sy, &S.sz, &S.nx, &S.ny, &S.nz})
    RVP_TRY(t->alloc(fb, S.s, true));
  for (Tmp *t : {&S.aos, &S.vox, &S.sor, &S.naos}) RVP_TRY(t->alloc(3 * fb, S.s, true));
  for (Tmp *t : {&S.off, &S.voff, &S.soff, &S.status}) RVP_TRY(t->alloc((max_frames + 1) * sizeof(uint32_t), S.s, true));
  RVP_TRY(S.leaf.alloc(max_frames * sizeof(float), S.s, true));
  RVP_TRY(S.radius.alloc(max_frames * sizeof(float), S.s, true));
  RVP_HIP(hipHostMalloc((void **)&S.h_off, 4 * (max_frames + 1) * sizeof(uint32_t), hipHostMallocDefault));
  return RVP_OK;
}

static void slot_free(Slot &S) {
  // the slot's stream is drained before any of its buffers is released (a batch that ended in an error can
  // leave work queued on it)
  if (S.s) (void)hipStreamSynchronize(S.s);
  for (Tmp *t : {&S.aos, &S.x, &S.y, &S.z, &S.vox, &S.vx, &S.vy, &S.vz, &S.sor, &S.sx, &S.sy, &S.sz, &S.nx, &S.ny,
                 &S.nz, &S.naos, &S.off, &S.voff, &S.soff, &S.status, &S.leaf, &S.radius})
    t->release();
  if (S.h_off) (void)hipHostFree(S.h_off);
  if (S.s) (void)rvp_stream_destroy(S.s);
}

// SoA normals -> AoS for the first min(*n_dev, n_bound) points (the count a stage kept, read on the device)
__global__ void soa_to_aos_bounded_k(const float *__restrict__ x, const float *__restrict__ y,
                                     const float *__restrict__ z, size_t n_bound, const uint32_t *__restrict__ n_dev,
                                     float *__restrict__ aos) {
  const size_t n = min((size_t)*n_dev, n_bound);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    aos[3 * i] = x[i];
    aos[3 * i + 1] = y[i];
    aos[3 * i + 2] = z[i];
  }
}

// one batch through the pipeline; fills every request's results
static int run_batch(const rvp_frame_params &P, Slot &S, std::vector<Request *> &batch) {
  const int ncl = (int)batch.size();
  hipStream_t s = S.s;
  uint32_t *h_off = S.h_off;
  size_t total = 0;
  h_off[0] = 0;
  // each frame goes up straight from its pinned copy
  for (int c = 0; c < ncl; ++c) {
    const size_t n = batch[c]->n;
    if (n) RVP_HIP(hipMemcpyAsync(S.aos.as<float>() + 3 * total, batch[c]->xyz, n * 3 * sizeof(float),
                                  hipMemcpyHostToDevice, s));
    total += n;
    h_off[c + 1] = (uint32_t)total;
  }
  const size_t nb = std::max<size_t>(total, 1);
  float *h_par = reinterpret_cast<float *>(h_off + 2 * (ncl + 1));
  for (int c = 0; c < ncl; ++c) {
    h_par[c] = P.leaf;
    h_par[ncl + c] = P.normal_radius;
  }
  RVP_HIP(hipMemcpyAsync(S.off.p, h_off, (ncl + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  RVP_HIP(hipMemcpyAsync(S.leaf.p, h_par, ncl * sizeof(float), hipMemcpyHostToDevice, s));
  RVP_HIP(hipMemcpyAsync(S.radius.p, h_par + ncl, ncl * sizeof(float), hipMemcpyHostToDevice, s));
  float *x = S.x.as<float>(), *y = S.y.as<float>(), *z = S.z.as<float>();
  RVP_TRY(rvp_aos_to_soa(S.aos.as<float>(), total, x, y, z, s));
  RVP_TRY(rvp_batch_voxel(x, y, z, S.off.as<uint32_t>(), ncl, nb, S.leaf.as<float>(), S.vox.as<float>(),
                          S.voff.as<uint32_t>(), S.status.as<uint32_t>(), s));
  RVP_TRY(rvp_aos_to_soa_bounded(S.vox.as<float>(), nb, S.voff.as<uint32_t>() + ncl, S.vx.as<float>(),
                                 S.vy.as<float>(), S.vz.as<float>(), s));
  RVP_TRY(rvp_batch_sor_exact(S.vx.as<float>(), S.vy.as<float>(), S.vz.as<float>(), S.voff.as<uint32_t>(), ncl, nb,
                              P.sor_k, P.sor_alpha, S.sor.as<float>(), S.soff.as<uint32_t>(), nullptr, nullptr, s));
  float *sx = S.sx.as<float>(), *sy = S.sy.as<float>(), *sz = S.sz.as<float>();
  RVP_TRY(rvp_aos_to_soa_bounded(S.sor.as<float>(), nb, S.soff.as<uint32_t>() + ncl, sx, sy, sz, s));
  RVP_TRY(rvp_batch_normals_radius(sx, sy, sz, S.soff.as<uint32_t>(), ncl, nb, S.radius.as<float>(), nullptr,
                                   P.normal_max_nn, P.vp[0], P.vp[1], P.vp[2], S.nx.as<float>(), S.ny.as<float>(),
                                   S.nz.as<float>(), s));
  // the normals interleaved on the device, so each frame's results come back in two plain copies
  RVP_LAUNCH("soa_to_aos_bounded_k", s, soa_to_aos_bounded_k<<<(unsigned)std::min<size_t>(1024, ceil_div(nb, 256)),
                                                               256, 0, s>>>(S.nx.as<float>(), S.ny.as<float>(),
                                                                            S.nz.as<float>(), nb,
                                                                            S.soff.as<uint32_t>() + ncl,
                                                                            S.naos.as<float>()));
  std::vector<float> models(4 * ncl);
  std::vector<int> inl(ncl), its(ncl);
  RVP_TRY(rvp_batch_ransac_plane(sx, sy, sz, S.soff.as<uint32_t>(), ncl, nb, P.ransac_thresh, P.ransac_iters, 1e-6f,
                                 P.ransac_prob, models.data(), inl.data(), its.data(), s));
  // per-frame counts, then each frame's kept points and normals straight into its pinned result buffers
  uint32_t *h_voff = h_off + (ncl + 1);
  RVP_HIP(hipMemcpyAsync(h_voff, S.voff.p, (ncl + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  std::vector<uint32_t> h_soff(ncl + 1), h_status(ncl);
  RVP_HIP(hipMemcpyAsync(h_soff.data(), S.soff.p, (ncl + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipMemcpyAsync(h_status.data(), S.status.p, ncl * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < ncl; ++c) {
    rvp_frame_result &r = batch[c]->res;
    const size_t a = h_soff[c], m = h_soff[c + 1] - h_soff[c];
    r.points = g_pinned.get(std::max<size_t>(m, 1) * 3);
    r.normals = g_pinned.get(std::max<size_t>(m, 1) * 3);
    if (!r.points || !r.normals) return fail(RVP_E_OOM, "pinned result allocation");
    if (m) {
      RVP_HIP(hipMemcpyAsync(r.points, S.sor.as<float>() + 3 * a, m * 3 * sizeof(float), hipMemcpyDeviceToHost, s));
      RVP_HIP(hipMemcpyAsync(r.normals, S.naos.as<float>() + 3 * a, m * 3 * sizeof(float), hipMemcpyDeviceToHost,
                             s));
    }
  }
  RVP_HIP(hipStreamSynchronize(s));
  for (int c = 0; c < ncl; ++c) {
    rvp_frame_result &r = batch[c]->res;
    r.n_in = h_off[c + 1] - h_off[c];
    r.n_voxel = h_voff[c + 1] - h_voff[c];
    r.n_sor = h_soff[c + 1] - h_soff[c];
    r.voxel_status = h_status[c];
    for (int k = 0; k < 4; ++k) r.model[k] = models[4 * c + k];
    r.inliers = inl[c];
    r.iterations = its[c];
    r.batch_frames = ncl;
  }
  return RVP_OK;
}

static void worker(rvp_server *srv, Slot *S) {
  (void)hipSetDevice(S->dev);
  for (;;) {
    std::vector<Request *> batch;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv_work.wait(lk, [&] { return srv->stopping || !srv->queue.empty(); });
      if (srv->queue.empty()) return;  // stopping with nothing left
      size_t pts = 0;
      while (!srv->queue.empty() && (int)batch.size() < srv->max_frames) {
        Request *r = srv->queue.front();
        const size_t n = r->n;
        if (!batch.empty() && pts + n > S->cap) break;
        srv->queue.pop_front();
        batch.push_back(r);
        pts += n;
      }
    }
    const double t0 = now_s();
    for (Request *r : batch) r->res.t_start = t0;
    const int st = run_batch(srv->p, *S, batch);
    const std::string err = st == RVP_OK ? std::string() : std::string(rvp_last_error());
    const double t1 = now_s();
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      for (Request *r : batch) {
        r->res.t_done = t1;
        r->status = st;
        r->error = err;
        r->done = true;
      }
    }
    srv->cv_done.notify_all();
  }
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_server_create(rvp_server **out, const rvp_frame_params *params, int slots, size_t slot_points,
                      size_t hbm_budget) {
  RVP_CHECK(out && params, RVP_E_ARG, "null argument");
  RVP_CHECK(slots >= 1 && slots <= 16, RVP_E_ARG, "slots must be in [1, 16]");
  RVP_CHECK(slot_points >= 1024 && slot_points < 0x7fffffffull / 3, RVP_E_ARG, "slot_points out of range");
  RVP_CHECK(params->sor_k >= 1 && params->sor_k + 1 <= MAX_TOPK_API, RVP_E_UNSUPPORTED, "sor_k must be in [1, 63]");
  RVP_CHECK(params->normal_max_nn >= 1 && params->normal_max_nn <= MAX_TOPK_API, RVP_E_UNSUPPORTED,
            "normal_max_nn must be in [1, 64]");
  const size_t per_slot = slot_points * SERVER_BYTES_PER_POINT;
  RVP_CHECK(hbm_budget >= per_slot, RVP_E_OOM, "HBM budget below one slot's working set");
  // admission by budget: as many slots as the budget holds, at most the number asked for
  slots = (int)std::min<size_t>((size_t)slots, hbm_budget / per_slot);
  size_t free_b = 0, total_b = 0;
  RVP_HIP(hipMemGetInfo(&free_b, &total_b));
  RVP_CHECK((size_t)slots * per_slot <= free_b, RVP_E_OOM, "device memory below the slots' working sets");
  int dev = 0;
  RVP_HIP(hipGetDevice(&dev));
  rvp_server *srv = new rvp_server();
  srv->p = *params;
  srv->slot_points = slot_points;
  for (int i = 0; i < slots; ++i) {
    Slot *S = new Slot();
    const int st = slot_alloc(*S, dev, slot_points, srv->max_frames);
    if (st != RVP_OK) {
      slot_free(*S);
      delete S;
      for (Slot *o : srv->slots) {
        slot_free(*o);
        delete o;
      }
      delete srv;
      return st;
    }
    srv->slots.push_back(S);
  }
  for (Slot *S : srv->slots) S->th = std::thread(worker, srv, S);
  *out = srv;
  return RVP_OK;
}

int rvp_server_slots(const rvp_server *srv, int *slots) {
  RVP_CHECK(srv && slots, RVP_E_ARG, "null argument");
  *slots = (int)srv->slots.size();
  return RVP_OK;
}

int rvp_server_submit(rvp_server *srv, const float *xyz, size_t n, uint64_t *ticket) {
  RVP_CHECK(srv && ticket && (xyz || n == 0), RVP_E_ARG, "null argument");
  RVP_CHECK(n <= srv->slot_points, RVP_E_UNSUPPORTED, "frame larger than a slot");
  Request *r = new Request();
  r->n = n;
  r->xyz = g_pinned.get(std::max<size_t>(n, 1) * 3);
  if (!r->xyz) {
    delete r;
    return fail(RVP_E_OOM, "pinned staging allocation");
  }
  if (n) std::memcpy(r->xyz, xyz, n * 3 * sizeof(float));
  r->res.t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    RVP_CHECK(!srv->stopping, RVP_E_STATE, "server is stopping");
    r->ticket = srv->next_ticket++;
    r->res.ticket = r->ticket;
    srv->live[r->ticket] = r;
    srv->queue.push_back(r);
    *ticket = r->ticket;
  }
  srv->cv_work.notify_one();
  return RVP_OK;
}

int rvp_server_wait(rvp_server *srv, uint64_t ticket, rvp_frame_result *out) {
  RVP_CHECK(srv && out, RVP_E_ARG, "null argument");
  Request *r = nullptr;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    RVP_CHECK(it != srv->live.end(), RVP_E_ARG, "unknown ticket");
    r = it->second;
    srv->cv_done.wait(lk, [&] { return r->done; });
    srv->live.erase(it);
  }
  const int st = r->status;
  const std::string err = r->error;
  *out = r->res;  // ownership of points / normals passes to the caller
  delete r;
  if (st != RVP_OK) {
    g_pinned.put(out->points);
    g_pinned.put(out->normals);
    out->points = out->normals = nullptr;
    return fail(st, "batch failed: " + err);
  }
  return RVP_OK;
}

int rvp_result_free(rvp_frame_result *r) {
  RVP_CHECK(r, RVP_E_ARG, "null argument");
  g_pinned.put(r->points);
  g_pinned.put(r->normals);
  r->points = r->normals = nullptr;
  return RVP_OK;
}

int rvp_server_destroy(rvp_server *srv) {
  if (!srv) return RVP_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stopping = true;
  }
  srv->cv_work.notify_all();
  for (Slot *S : srv->slots)
    if (S->th.joinable()) S->th.join();
  for (Slot *S : srv->slots) {
    slot_free(*S);
    delete S;
  }
  for (auto &kv : srv->live) {
    g_pinned.put(kv.second->res.points);
    g_pinned.put(kv.second->res.normals);
    delete kv.second;
  }
  delete srv;
  return RVP_OK;
}

}  // extern "C"
