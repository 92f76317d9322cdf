// cluster.hip -- Euclidean clustering (euclidean_clustering.cpp:113-189) as the connected components of the
// radius graph.
//
// The reference grows each cluster by breadth-first search from the smallest unvisited index; its clusters
// are exactly the connected components of the graph that links points within the tolerance
// (D1 <= fl(tol * tol), symmetric in the two points), listed by smallest index, each sorted.  Here one wave
// per point walks the point's radius window on the hash grid and unites it with every hit of smaller index
// in a lock-free union-find whose roots are always the smallest index of their set (a root is only ever
// linked under a smaller root; path halving only moves pointers to smaller indices, so no cycle can form).
// A stable radix sort of (root, index) then lists the clusters by root with ascending indices, and the size
// filter compacts them.
#include <algorithm>

#include "prims.hpp"
#include "query.hpp"

namespace rvp {

constexpr uint32_t UF_LIMIT = 1u << 22;  // steps after which a find / unite gives up and counts an error
__device__ unsigned long long g_uf_errors;

__device__ __forceinline__ uint32_t uf_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ uint32_t uf_find(uint32_t *parent, uint32_t x) {
  for (uint32_t it = 0; it < UF_LIMIT; ++it) {
    const uint32_t p = uf_load(parent + x);
    if (p == x) return x;
    const uint32_t gp = uf_load(parent + p);
    if (gp != p) atomicCAS(parent + x, p, gp);  // path halving
    x = gp;
  }
  atomicAdd(&g_uf_errors, 1ull);
  return x;
}

__device__ void uf_unite(uint32_t *parent, uint32_t a, uint32_t b) {
  for (uint32_t it = 0; it < UF_LIMIT; ++it) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS(parent + a, a, b) == a) return;  // the larger root goes under the smaller
  }
  atomicAdd(&g_uf_errors, 1ull);
}

__global__ void uf_init_k(uint32_t n, uint32_t *__restrict__ parent) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = i;
}

struct UniteSink {
  uint32_t *parent;
  uint32_t qi;
  __device__ __forceinline__ void operator()(bool, bool hit, uint64_t key, uint32_t) {
    if (hit) {
      const uint32_t j = key_idx(key);
      if (j < qi) uf_unite(parent, qi, j);
    }
  }
};

// one wave per point (cell order): unite with every neighbour of smaller index within the tolerance
__global__ __launch_bounds__(QT) void cluster_unite_k(GridView g, float tol, uint32_t *__restrict__ parent) {
  __shared__ WaveLds lds[QW];
  const int w = wave_in_block();
  const float r2 = __fmul_rn(tol, tol);
  for (uint32_t q = blockIdx.x * QW + w; q < g.n; q += gridDim.x * QW) {
    const float x = g.xs[q], y = g.ys[q], z = g.zs[q];
    int lo[3], hi[3];
    window3(g, x, y, z, tol, lo, hi);
    UniteSink sink{parent, g.perm[q]};
    scan_box(g, x, y, z, lo, hi, false, lo, hi, r2, true, lds[w], sink);
  }
}

// label = root (the component's smallest index); sizes counted once per run of equal labels in a wave
__global__ void cluster_label_k(uint32_t n, uint32_t *__restrict__ parent, uint32_t *__restrict__ label,
                                uint32_t *__restrict__ size) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = i < n;
  uint32_t r = 0;
  if (ok) {
    r = uf_find(parent, i);
    label[i] = r;
  }
  uint64_t pending = ballot(ok);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lr = readlaneu(r, leader);
    const uint64_t same = ballot(ok && r == lr) & pending;
    if (lane() == leader) atomicAdd(&size[lr], (uint32_t)popc(same));
    pending &= ~same;
  }
}

// over the (label, index) pairs sorted by label: kept entries and the first entry of each kept cluster
__global__ void cluster_flags_k(uint32_t n, const uint32_t *__restrict__ lab, const uint32_t *__restrict__ size,
                                int min_size, int max_size, uint32_t *__restrict__ keep,
                                uint32_t *__restrict__ first) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t s = size[lab[j]];
  const bool k = (int64_t)s >= (int64_t)min_size && (int64_t)s <= (int64_t)max_size;
  keep[j] = k ? 1u : 0u;
  first[j] = (k && (j == 0 || lab[j - 1] != lab[j])) ? 1u : 0u;
}

__global__ void cluster_emit_k(uint32_t n, const uint32_t *__restrict__ idx, const uint32_t *__restrict__ keep,
                               const uint32_t *__restrict__ kpos, const uint32_t *__restrict__ first,
                               const uint32_t *__restrict__ cpos, const uint32_t *__restrict__ kept_total,
                               int32_t *__restrict__ out_idx, int64_t *__restrict__ out_off) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0) out_off[0] = 0;
  if (j >= n) return;
  if (keep[j]) out_idx[kpos[j]] = (int32_t)idx[j];
  // cluster c ends where cluster c + 1 starts; the last one at the kept total
  if (first[j] && cpos[j] > 0) out_off[cpos[j]] = (int64_t)kpos[j];
  if (j == n - 1) out_off[cpos[j] + first[j]] = (int64_t)*kept_total;
}

}  // namespace rvp

using namespace rvp;

extern "C" {

int rvp_euclidean_cluster(const rvp_hash *h, float tolerance, int min_size, int max_size, int32_t *indices_out,
                          int64_t *offsets_out, size_t *nclusters, int32_t *labels, rvp_stream_t st) {
  RVP_CHECK(h && nclusters, RVP_E_ARG, "null argument");
  RVP_CHECK(tolerance >= 0.0f, RVP_E_ARG, "negative tolerance");
  *nclusters = 0;
  const Grid &G = h->g;
  RVP_CHECK(G.built, RVP_E_STATE, "hash not built");
  const size_t n = G.n;
  RVP_CHECK(offsets_out && (indices_out || n == 0), RVP_E_ARG, "null output");
  RVP_CHECK(n < 0x7fffffffull, RVP_E_UNSUPPORTED, "more than 2^31-1 points");
  RVP_TRY(pool_init());
  hipStream_t s = S(st);
  if (n == 0) {
    const int64_t z = 0;
    RVP_HIP(hipMemcpyAsync(offsets_out, &z, sizeof(z), hipMemcpyHostToDevice, s));
    RVP_HIP(hipStreamSynchronize(s));
    return RVP_OK;
  }
  const unsigned nb = ceil_div(n, 256);
  Tmp parent, label, size, vals, kt, vt, keep, kpos, first, cpos, totals;
  RVP_TRY(parent.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(label.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(size.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(totals.alloc(2 * sizeof(uint32_t), s));
  RVP_HIP(hipMemsetAsync(size.p, 0, n * sizeof(uint32_t), s));
  RVP_LAUNCH("uf_init_k", s, uf_init_k<<<nb, 256, 0, s>>>((uint32_t)n, parent.as<uint32_t>()));
  const unsigned qb = (unsigned)std::min<size_t>(16384, ceil_div(n, QW));
  RVP_LAUNCH("cluster_unite_k", s, cluster_unite_k<<<qb, QT, 0, s>>>(G.view(), tolerance, parent.as<uint32_t>()));
  RVP_LAUNCH("cluster_label_k", s, cluster_label_k<<<nb, 256, 0, s>>>((uint32_t)n, parent.as<uint32_t>(),
                                                                      label.as<uint32_t>(), size.as<uint32_t>()));
  if (labels) RVP_HIP(hipMemcpyAsync(labels, label.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  // (label, index) pairs sorted by label, stable: indices ascend inside each cluster
  RVP_TRY(vals.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(kt.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(vt.alloc(n * sizeof(uint32_t), s));
  RVP_LAUNCH("uf_init_k", s, uf_init_k<<<nb, 256, 0, s>>>((uint32_t)n, vals.as<uint32_t>()));
  uint32_t *kp = label.as<uint32_t>(), *vp = vals.as<uint32_t>(), *ktp = kt.as<uint32_t>(), *vtp = vt.as<uint32_t>();
  RVP_TRY(radix_sort_pairs_u32(&kp, &vp, &ktp, &vtp, n, bits_for(n - 1), s, nullptr, nullptr, n >= (1u << 19)));
  RVP_TRY(keep.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(kpos.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(first.alloc(n * sizeof(uint32_t), s));
  RVP_TRY(cpos.alloc(n * sizeof(uint32_t), s));
  RVP_LAUNCH("cluster_flags_k", s, cluster_flags_k<<<nb, 256, 0, s>>>((uint32_t)n, kp, size.as<uint32_t>(), min_size,
                                                                      max_size, keep.as<uint32_t>(),
                                                                      first.as<uint32_t>()));
  RVP_TRY(exclusive_scan_u32(keep.as<uint32_t>(), kpos.as<uint32_t>(), n, totals.as<uint32_t>(), s));
  RVP_TRY(exclusive_scan_u32(first.as<uint32_t>(), cpos.as<uint32_t>(), n, totals.as<uint32_t>() + 1, s));
  RVP_LAUNCH("cluster_emit_k", s, cluster_emit_k<<<nb, 256, 0, s>>>((uint32_t)n, vp, keep.as<uint32_t>(),
                                                                    kpos.as<uint32_t>(), first.as<uint32_t>(),
                                                                    cpos.as<uint32_t>(), totals.as<uint32_t>(),
                                                                    indices_out, offsets_out));
  uint32_t tot[2] = {0, 0};
  RVP_HIP(hipMemcpyAsync(tot, totals.p, sizeof(tot), hipMemcpyDeviceToHost, s));
  RVP_HIP(hipStreamSynchronize(s));
  *nclusters = tot[1];
  return RVP_OK;
}

int rvp_cluster_errors(unsigned long long *count) {
  RVP_CHECK(count, RVP_E_ARG, "null argument");
  RVP_HIP(hipDeviceSynchronize());
  RVP_HIP(hipMemcpyFromSymbol(count, HIP_SYMBOL(g_uf_errors), sizeof(*count)));
  return RVP_OK;
}

}  // extern "C"
