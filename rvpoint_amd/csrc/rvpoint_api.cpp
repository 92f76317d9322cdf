// rvpoint_api.cpp -- the reference's rvpoint:: API (the "rvv_pcl.h" signatures) implemented over the C ABI
// of librvp_hip.so.  Host buffers are staged to HBM, the GPU computes, results are copied back; there is no
// CPU compute path.  Device errors throw rvpoint::DeviceError (the reference has no error channel).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rvp.h"
#include "rvv_pcl.h"

namespace rvpoint {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

namespace {

void check(int st, const char *what) {
  if (st != RVP_OK) throw DeviceError(std::string(what) + ": " + rvp_last_error());
}

// RAII device buffer
struct Dev {
  void *p = nullptr;
  Dev() = default;
  explicit Dev(std::size_t bytes) { check(rvp_malloc(&p, bytes ? bytes : 16), "rvp_malloc"); }
  Dev(const Dev &) = delete;
  Dev &operator=(const Dev &) = delete;
  Dev(Dev &&o) noexcept : p(o.p) { o.p = nullptr; }
  Dev &operator=(Dev &&o) noexcept {
    if (this != &o) {
      if (p) rvp_free(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~Dev() {
    if (p) rvp_free(p);
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

Dev upload(const void *src, std::size_t bytes) {
  Dev d(bytes);
  if (bytes) check(rvp_memcpy_h2d(d.p, src, bytes, nullptr), "rvp_memcpy_h2d");
  return d;
}

void download(void *dst, const void *src, std::size_t bytes) {
  if (bytes) check(rvp_memcpy_d2h(dst, src, bytes, nullptr), "rvp_memcpy_d2h");
}

struct DevCloud {
  Dev x, y, z;
  std::size_t n = 0;
};

DevCloud upload_soa(const float *x, const float *y, const float *z, std::size_t n) {
  DevCloud c;
  c.n = n;
  c.x = upload(x, n * sizeof(float));
  c.y = upload(y, n * sizeof(float));
  c.z = upload(z, n * sizeof(float));
  return c;
}

DevCloud upload_aos(const PointXYZ *in, std::size_t n) {
  std::vector<float> x(n), y(n), z(n);
  for (std::size_t i = 0; i < n; ++i) {
    x[i] = in[i].x;
    y[i] = in[i].y;
    z[i] = in[i].z;
  }
  return upload_soa(x.data(), y.data(), z.data(), n);
}

}  // namespace

// ---------------------------------------------------------------------------------------------------
// D1 / D2
// ---------------------------------------------------------------------------------------------------
void get_dist_sq_rvv(const float *x, const float *y, const float *z, float qx, float qy, float qz, float *out_d2,
                     std::size_t n) {
  if (n == 0) return;
  DevCloud c = upload_soa(x, y, z, n);
  Dev o(n * sizeof(float));
  check(rvp_dist_sq(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), n, qx, qy, qz, o.as<float>(), nullptr),
        "rvp_dist_sq");
  download(out_d2, o.p, n * sizeof(float));
}

void get_inds_in_radius_rvv(const float *x, const float *y, const float *z, const int *subset, std::size_t n,
                            float qx, float qy, float qz, float r2, std::vector<int> &out_indices,
                            std::vector<float> &out_dists) {
  if (n == 0) return;
  // the subset indexes the full arrays, whose length the reference does not pass: upload the gathered
  // subset coordinates and index them 0..n-1
  std::vector<float> gx(n), gy(n), gz(n);
  std::vector<int> local(n);
  for (std::size_t i = 0; i < n; ++i) {
    gx[i] = x[subset[i]];
    gy[i] = y[subset[i]];
    gz[i] = z[subset[i]];
    local[i] = (int)i;
  }
  DevCloud c = upload_soa(gx.data(), gy.data(), gz.data(), n);
  Dev sub = upload(local.data(), n * sizeof(int));
  Dev oi(n * sizeof(int)), od(n * sizeof(float));
  std::size_t m = 0;
  check(rvp_inds_in_radius(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), sub.as<int32_t>(), n, qx, qy, qz, r2,
                           oi.as<int32_t>(), od.as<float>(), &m, nullptr),
        "rvp_inds_in_radius");
  std::vector<int> li(m);
  std::vector<float> ld(m);
  download(li.data(), oi.p, m * sizeof(int));
  download(ld.data(), od.p, m * sizeof(float));
  for (std::size_t i = 0; i < m; ++i) {
    out_indices.push_back(subset[li[i]]);
    out_dists.push_back(ld[i]);
  }
}

// ---------------------------------------------------------------------------------------------------
// SpatialHash
// ---------------------------------------------------------------------------------------------------
SpatialHash::SpatialHash() { check(rvp_hash_create(&h_), "rvp_hash_create"); }

SpatialHash::~SpatialHash() {
  rvp_hash_destroy(h_);
  rvp_free(dx_);
  rvp_free(dy_);
  rvp_free(dz_);
}

void SpatialHash::setInputCloud(const PointCloudSoA &cloud, float cell_size) {
  cloud_ = cloud;
  cell_size_ = cell_size;
}

void SpatialHash::build() {
  const std::size_t n = cloud_.n;
  if (n > dev_n_ || !dx_) {
    rvp_free(dx_);
    rvp_free(dy_);
    rvp_free(dz_);
    dx_ = dy_ = dz_ = nullptr;
    void *p;
    check(rvp_malloc(&p, (n ? n : 1) * sizeof(float)), "rvp_malloc"); dx_ = (float *)p;
    check(rvp_malloc(&p, (n ? n : 1) * sizeof(float)), "rvp_malloc"); dy_ = (float *)p;
    check(rvp_malloc(&p, (n ? n : 1) * sizeof(float)), "rvp_malloc"); dz_ = (float *)p;
    dev_n_ = n;
  }
  if (n) {
    check(rvp_memcpy_h2d(dx_, cloud_.x, n * sizeof(float), nullptr), "h2d");
    check(rvp_memcpy_h2d(dy_, cloud_.y, n * sizeof(float), nullptr), "h2d");
    check(rvp_memcpy_h2d(dz_, cloud_.z, n * sizeof(float), nullptr), "h2d");
  }
  check(rvp_hash_build(h_, dx_, dy_, dz_, n, cell_size_, nullptr), "rvp_hash_build");
}

std::size_t SpatialHash::radiusSearch(const PointXYZ &query, float radius, std::vector<int> &indices,
                                      std::vector<float> &dists, int max_nn) const {
  indices.clear();
  dists.clear();
  const float q[3] = {query.x, query.y, query.z};
  Dev dq = upload(q, sizeof(q));
  const float *qx = dq.as<float>(), *qy = qx + 1, *qz = qx + 2;
  if (max_nn > 0) {
    Dev oi((std::size_t)max_nn * sizeof(int)), od((std::size_t)max_nn * sizeof(float)), oc(sizeof(int));
    check(rvp_radius_search(h_, qx, qy, qz, 1, radius, max_nn, oi.as<int32_t>(), od.as<float>(), oc.as<int32_t>(),
                            nullptr, nullptr),
          "rvp_radius_search");
    int m = 0;
    download(&m, oc.p, sizeof(int));
    indices.resize(m);
    dists.resize(m);
    download(indices.data(), oi.p, m * sizeof(int));
    download(dists.data(), od.p, m * sizeof(float));
    return indices.size();
  }
  Dev cnt(sizeof(int)), off(2 * sizeof(int64_t));
  check(rvp_radius_count(h_, qx, qy, qz, 1, radius, cnt.as<int32_t>(), nullptr), "rvp_radius_count");
  int64_t tot = 0;
  check(rvp_exclusive_scan(cnt.as<int32_t>(), 1, off.as<int64_t>(), &tot, nullptr), "rvp_exclusive_scan");
  Dev oi((std::size_t)tot * sizeof(int)), od((std::size_t)tot * sizeof(float));
  check(rvp_radius_fill(h_, qx, qy, qz, 1, radius, off.as<int64_t>(), oi.as<int32_t>(), od.as<float>(), nullptr),
        "rvp_radius_fill");
  indices.resize(tot);
  dists.resize(tot);
  download(indices.data(), oi.p, tot * sizeof(int));
  download(dists.data(), od.p, tot * sizeof(float));
  return indices.size();
}

void SpatialHash::radiusSearchAll(float radius, int max_nn, std::vector<int> &indices, std::vector<float> &dists,
                                  std::vector<int> &counts) const {
  const std::size_t n = cloud_.n;
  if (max_nn <= 0) throw std::invalid_argument("radiusSearchAll needs max_nn > 0");
  indices.assign(n * max_nn, -1);
  dists.assign(n * max_nn, 0.0f);
  counts.assign(n, 0);
  if (n == 0) return;
  Dev oi(n * max_nn * sizeof(int)), od(n * max_nn * sizeof(float)), oc(n * sizeof(int));
  check(rvp_radius_search(h_, nullptr, nullptr, nullptr, n, radius, max_nn, oi.as<int32_t>(), od.as<float>(),
                          oc.as<int32_t>(), nullptr, nullptr),
        "rvp_radius_search");
  download(indices.data(), oi.p, n * max_nn * sizeof(int));
  download(dists.data(), od.p, n * max_nn * sizeof(float));
  download(counts.data(), oc.p, n * sizeof(int));
}

// ---------------------------------------------------------------------------------------------------
// brute-force radius search
// ---------------------------------------------------------------------------------------------------
static std::size_t radius_bf(const DevCloud &c, PointXYZ q, float radius, int *indices, float *dists, int max_nn) {
  if (max_nn <= 0) return 0;  // radius_search.cpp:22-29 writes nothing and returns 0
  const float qq[3] = {q.x, q.y, q.z};
  Dev dq = upload(qq, sizeof(qq));
  Dev oi((std::size_t)max_nn * sizeof(int)), od((std::size_t)max_nn * sizeof(float)), oc(sizeof(int));
  const float *qx = dq.as<float>();
  check(rvp_radius_bruteforce(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), c.n, qx, qx + 1, qx + 2, 1, radius,
                              max_nn, oi.as<int32_t>(), od.as<float>(), oc.as<int32_t>(), nullptr),
        "rvp_radius_bruteforce");
  int m = 0;
  download(&m, oc.p, sizeof(int));
  download(indices, oi.p, m * sizeof(int));
  if (dists) download(dists, od.p, m * sizeof(float));
  return (std::size_t)m;
}

std::size_t radius_search_sc(const PointXYZ *cloud, std::size_t n, PointXYZ query, float radius, int *indices,
                             float *dists, int max_nn) {
  if (n == 0) return 0;
  return radius_bf(upload_aos(cloud, n), query, radius, indices, dists, max_nn);
}

std::size_t radius_search_rvv(const PointCloudSoA &cloud, PointXYZ query, float radius, int *indices, float *dists,
                              int max_nn) {
  if (cloud.n == 0) return 0;
  return radius_bf(upload_soa(cloud.x, cloud.y, cloud.z, cloud.n), query, radius, indices, dists, max_nn);
}

// ---------------------------------------------------------------------------------------------------
// voxel grid
// ---------------------------------------------------------------------------------------------------
static std::size_t voxel(const DevCloud &c, PointXYZ *out, float leaf) {
  Dev o(c.n * 3 * sizeof(float));
  std::size_t m = 0;
  check(rvp_voxel(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), c.n, leaf, o.as<float>(), &m, nullptr),
        "rvp_voxel");
  download(out, o.p, m * sizeof(PointXYZ));
  return m;
}

std::size_t voxel_grid_downsamp_sc(const PointXYZ *in, std::size_t n, PointXYZ *out, float leaf_size) {
  if (n == 0) return 0;
  return voxel(upload_aos(in, n), out, leaf_size);
}

std::size_t voxel_grid_downsamp_rvv(const PointCloudSoA &in, PointXYZ *out, float leaf_size) {
  if (in.n == 0) return 0;
  return voxel(upload_soa(in.x, in.y, in.z, in.n), out, leaf_size);
}

std::size_t voxel_grid_downsamp_rvv_v2(const PointCloudSoA &in, PointXYZ *out, float leaf_size) {
  return voxel_grid_downsamp_rvv(in, out, leaf_size);
}

// ---------------------------------------------------------------------------------------------------
// SOR
// ---------------------------------------------------------------------------------------------------
static std::size_t sor_exact(const DevCloud &c, PointXYZ *out, int k, float alpha) {
  Dev o(c.n * 3 * sizeof(float));
  std::size_t m = 0;
  check(rvp_sor_exact(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), c.n, k, alpha, o.as<float>(), &m, nullptr,
                      nullptr, nullptr),
        "rvp_sor_exact");
  download(out, o.p, m * sizeof(PointXYZ));
  return m;
}

std::size_t sor_sc(const PointXYZ *in, std::size_t n, PointXYZ *out, int k, float alpha) {
  if (n == 0) return 0;
  return sor_exact(upload_aos(in, n), out, k, alpha);
}

std::size_t sor_rvv(const PointCloudSoA &in, PointXYZ *out, int k, float alpha) {
  if (in.n == 0) return 0;
  return sor_exact(upload_soa(in.x, in.y, in.z, in.n), out, k, alpha);
}

std::size_t sor_spatial_hash(const PointCloudSoA &in, const SpatialHash &hash, PointXYZ *out, int k, float alpha,
                             float search_radius) {
  if (in.n == 0) return 0;
  if (hash.cloud().n != in.n) throw std::invalid_argument("sor_spatial_hash: hash not built on this cloud");
  Dev o(in.n * 3 * sizeof(float));
  std::size_t m = 0;
  check(rvp_sor_radius(hash.handle(), k, alpha, search_radius, o.as<float>(), &m, nullptr, nullptr, nullptr),
        "rvp_sor_radius");
  download(out, o.p, m * sizeof(PointXYZ));
  return m;
}

// ---------------------------------------------------------------------------------------------------
// normals
// ---------------------------------------------------------------------------------------------------
void normal_estimation_sc(const PointXYZ *in, std::size_t n, float *nx, float *ny, float *nz, int k, float,
                          float vp_x, float vp_y, float vp_z, int) {
  if (n == 0) return;
  DevCloud c = upload_aos(in, n);
  Dev ox(n * sizeof(float)), oy(n * sizeof(float)), oz(n * sizeof(float));
  check(rvp_normals_knn(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), n, k, vp_x, vp_y, vp_z, ox.as<float>(),
                        oy.as<float>(), oz.as<float>(), nullptr),
        "rvp_normals_knn");
  download(nx, ox.p, n * sizeof(float));
  download(ny, oy.p, n * sizeof(float));
  download(nz, oz.p, n * sizeof(float));
}

void normal_estimation_rvv_capped(const PointCloudSoA &in, const SpatialHash &hash, float *nx, float *ny, float *nz,
                                  float radius, int max_nn, float vp_x, float vp_y, float vp_z) {
  if (in.n == 0) return;
  if (hash.cloud().n != in.n) throw std::invalid_argument("normal_estimation_rvv: hash not built on this cloud");
  const std::size_t n = in.n;
  Dev ox(n * sizeof(float)), oy(n * sizeof(float)), oz(n * sizeof(float));
  check(rvp_normals_radius(hash.handle(), radius, max_nn, vp_x, vp_y, vp_z, ox.as<float>(), oy.as<float>(),
                           oz.as<float>(), nullptr),
        "rvp_normals_radius");
  download(nx, ox.p, n * sizeof(float));
  download(ny, oy.p, n * sizeof(float));
  download(nz, oz.p, n * sizeof(float));
}

void normal_estimation_rvv(const PointCloudSoA &in, const SpatialHash &hash, float *nx, float *ny, float *nz, int,
                           float radius, float vp_x, float vp_y, float vp_z, int) {
  normal_estimation_rvv_capped(in, hash, nx, ny, nz, radius, 0, vp_x, vp_y, vp_z);
}

// ---------------------------------------------------------------------------------------------------
// Octree / PointerOctree facades (octree.h:11-48, pointer_octree.h:8-58) over the device grid
// ---------------------------------------------------------------------------------------------------
namespace {

// A grid edge for searches of any radius: about `per_cell` points per cell over the dimensions the cloud
// actually spans (a plane is 2-D, a line 1-D).  Search results do not depend on it.
float facade_cell(const PointCloudSoA &c, int per_cell) {
  float mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = std::numeric_limits<float>::infinity();
    mx[a] = -std::numeric_limits<float>::infinity();
  }
  for (std::size_t i = 0; i < c.n; ++i) {
    const float v[3] = {c.x[i], c.y[i], c.z[i]};
    for (int a = 0; a < 3; ++a) {
      if (v[a] < mn[a]) mn[a] = v[a];
      if (v[a] > mx[a]) mx[a] = v[a];
    }
  }
  double ext[3], emax = 0.0;
  for (int a = 0; a < 3; ++a) {
    ext[a] = mx[a] >= mn[a] ? (double)mx[a] - (double)mn[a] : 0.0;
    emax = std::max(emax, ext[a]);
  }
  if (!(emax > 0.0) || !std::isfinite(emax)) return 1.0f;
  double vol = 1.0;
  int dims = 0;
  for (int a = 0; a < 3; ++a)
    if (ext[a] > emax * 1e-3) {
      vol *= ext[a];
      ++dims;
    }
  const double cell = std::pow(vol * (double)std::max(per_cell, 1) / (double)std::max<std::size_t>(c.n, 1),
                               1.0 / dims);
  return (float)std::max(cell, emax * 1e-6);
}

std::unique_ptr<SpatialHash> facade_build(const PointCloudSoA &cloud, int leaf_points) {
  auto h = std::make_unique<SpatialHash>();
  h->setInputCloud(cloud, facade_cell(cloud, std::max(1, leaf_points / 8)));
  h->build();
  return h;
}

}  // namespace

OctreeNode::~OctreeNode() {
  for (OctreeNode *c : children) delete c;
}

Octree::Octree() = default;
Octree::~Octree() = default;

void Octree::setInputCloud(const PointCloudSoA &cloud) { cloud_ = cloud; }

void Octree::build() { hash_ = facade_build(cloud_, max_points_per_leaf_); }

std::size_t Octree::radiusSearch(const PointXYZ &query, float radius, std::vector<int> &indices,
                                 std::vector<float> &dists, int) const {
  indices.clear();
  dists.clear();
  if (!hash_) return 0;  // octree.cpp:117: no tree, no neighbours
  return hash_->radiusSearch(query, radius, indices, dists, 0);
}

const SpatialHash &Octree::hash() const {
  if (!hash_) throw std::logic_error("Octree: build() has not been called");
  return *hash_;
}

PointerOctreeNode::PointerOctreeNode() = default;
PointerOctreeNode::~PointerOctreeNode() {
  for (PointerOctreeNode *c : children) delete c;
}

PointerOctree::PointerOctree() = default;
PointerOctree::~PointerOctree() = default;

void PointerOctree::setInputCloud(const PointCloudSoA &cloud) { cloud_ = cloud; }

void PointerOctree::build() {
  hash_ = facade_build(cloud_, max_points_per_leaf_);
  root_.reset();
}

std::size_t PointerOctree::radiusSearch(const PointXYZ &query, float radius, std::vector<int> &indices,
                                        std::vector<float> &dists) const {
  indices.clear();
  dists.clear();
  if (!hash_) return 0;
  return hash_->radiusSearch(query, radius, indices, dists, 0);
}

std::size_t PointerOctree::radiusSearchScalar(const PointXYZ &query, float radius, std::vector<int> &indices,
                                              std::vector<float> &dists) const {
  return radiusSearch(query, radius, indices, dists);
}

const PointerOctreeNode *PointerOctree::getRoot() const {
  if (!hash_) return nullptr;
  if (!root_) {
    auto r = std::make_unique<PointerOctreeNode>();
    const std::size_t n = cloud_.n;
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    if (n) {
      mn[0] = mx[0] = cloud_.x[0];
      mn[1] = mx[1] = cloud_.y[0];
      mn[2] = mx[2] = cloud_.z[0];
    }
    for (std::size_t i = 0; i < n; ++i) {
      const float v[3] = {cloud_.x[i], cloud_.y[i], cloud_.z[i]};
      for (int a = 0; a < 3; ++a) {
        if (v[a] < mn[a]) mn[a] = v[a];
        if (v[a] > mx[a]) mx[a] = v[a];
      }
    }
    r->min_x = mn[0];
    r->min_y = mn[1];
    r->min_z = mn[2];
    r->max_x = mx[0];
    r->max_y = mx[1];
    r->max_z = mx[2];
    r->indices.resize(n);
    for (std::size_t i = 0; i < n; ++i) r->indices[i] = (int)i;
    r->leaf_x.assign(cloud_.x, cloud_.x + n);
    r->leaf_y.assign(cloud_.y, cloud_.y + n);
    r->leaf_z.assign(cloud_.z, cloud_.z + n);
    root_ = std::move(r);
  }
  return root_.get();
}

const SpatialHash &PointerOctree::hash() const {
  if (!hash_) throw std::logic_error("PointerOctree: build() has not been called");
  return *hash_;
}

// normal_estimation.cpp:155-187, 218-246: every neighbour within the radius (the trees ignore max_nn)
void normal_estimation_rvv(const PointCloudSoA &in, const Octree &octree, float *nx, float *ny, float *nz, int,
                           float radius, float vp_x, float vp_y, float vp_z, int) {
  if (in.n == 0) return;
  normal_estimation_rvv_capped(in, octree.hash(), nx, ny, nz, radius, 0, vp_x, vp_y, vp_z);
}

void normal_estimation_rvv(const PointCloudSoA &in, const PointerOctree &octree, float *nx, float *ny, float *nz,
                           int, float radius, float vp_x, float vp_y, float vp_z, int) {
  if (in.n == 0) return;
  normal_estimation_rvv_capped(in, octree.hash(), nx, ny, nz, radius, 0, vp_x, vp_y, vp_z);
}

// normal_estimation.cpp:248-255 builds a PointerOctree on the cloud; here a grid with the radius as its edge
void normal_estimation_rvv(const PointCloudSoA &in, float *nx, float *ny, float *nz, int, float radius,
                           float vp_x, float vp_y, float vp_z, int) {
  if (in.n == 0) return;
  SpatialHash h;
  h.setInputCloud(in, radius > 0.0f && std::isfinite(radius) ? radius : facade_cell(in, 8));
  h.build();
  normal_estimation_rvv_capped(in, h, nx, ny, nz, radius, 0, vp_x, vp_y, vp_z);
}

// statistical_outlier_removal.cpp:220-244: all neighbours within the radius, sorted, the k smallest after
// the first -- the same sum, in the same order, as the hash variant's capped (d2, idx) list
std::size_t sor_octree(const PointCloudSoA &in, const Octree &tree, PointXYZ *out, int k, float alpha,
                       float search_radius) {
  if (in.n == 0) return 0;
  return sor_spatial_hash(in, tree.hash(), out, k, alpha, search_radius);
}

// statistical_outlier_removal.cpp:272-301 sums the k smallest in nth_element's order, which the C++ library
// leaves unspecified; here they are summed ascending, as in sor_octree
std::size_t sor_pointer_octree(const PointCloudSoA &in, const PointerOctree &tree, PointXYZ *out, int k,
                               float alpha, float search_radius) {
  if (in.n == 0) return 0;
  return sor_spatial_hash(in, tree.hash(), out, k, alpha, search_radius);
}

// ---------------------------------------------------------------------------------------------------
// RANSAC + extraction
// ---------------------------------------------------------------------------------------------------
static int ransac(const DevCloud &c, float dist_thresh, int max_iters, float *model, float collinear,
                  float probability) {
  int best = 0, iters = 0;
  check(rvp_ransac_plane(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), c.n, dist_thresh, max_iters, collinear,
                         probability, model, &best, &iters, nullptr),
        "rvp_ransac_plane");
  // the reference reseeds and advances the C library's generator (ransac_plane.cpp:60, 69-71): leave it in
  // the same state (three draws per executed iteration)
  std::srand(0);
  for (long i = 0; i < 3L * iters; ++i) (void)std::rand();
  return best;
}

int ransac_plane_sc(const PointXYZ *cloud, std::size_t n, float dist_thresh, int max_iters, float *model,
                    float collinear_thresh, float probability) {
  if (n < 3) return 0;
  return ransac(upload_aos(cloud, n), dist_thresh, max_iters, model, collinear_thresh, probability);
}

int ransac_plane_rvv(const PointCloudSoA &cloud, float dist_thresh, int max_iters, float *model,
                     float collinear_thresh, float probability) {
  if (cloud.n < 3) return 0;
  return ransac(upload_soa(cloud.x, cloud.y, cloud.z, cloud.n), dist_thresh, max_iters, model, collinear_thresh,
                probability);
}

void extract_plane_inliers_outliers_rvv(const PointCloudSoA &cloud, const float *model, float dist_thresh,
                                        PointXYZ *inliers, PointXYZ *outliers, std::size_t &n_inliers,
                                        std::size_t &n_outliers) {
  n_inliers = n_outliers = 0;
  if (cloud.n == 0) return;
  DevCloud c = upload_soa(cloud.x, cloud.y, cloud.z, cloud.n);
  Dev a(cloud.n * 3 * sizeof(float)), b(cloud.n * 3 * sizeof(float));
  check(rvp_plane_extract(c.x.as<float>(), c.y.as<float>(), c.z.as<float>(), cloud.n, model, dist_thresh,
                          inliers ? a.as<float>() : nullptr, &n_inliers, outliers ? b.as<float>() : nullptr,
                          &n_outliers, nullptr),
        "rvp_plane_extract");
  if (inliers) download(inliers, a.p, n_inliers * sizeof(PointXYZ));
  if (outliers) download(outliers, b.p, n_outliers * sizeof(PointXYZ));
}

std::size_t extract_plane_inliers_rvv(const PointCloudSoA &cloud, const float *model, float dist_thresh,
                                      PointXYZ *inliers) {
  std::size_t ni = 0, no = 0;
  extract_plane_inliers_outliers_rvv(cloud, model, dist_thresh, inliers, nullptr, ni, no);
  return ni;
}

std::size_t extract_plane_outliers_rvv(const PointCloudSoA &cloud, const float *model, float dist_thresh,
                                       PointXYZ *outliers) {
  std::size_t ni = 0, no = 0;
  extract_plane_inliers_outliers_rvv(cloud, model, dist_thresh, nullptr, outliers, ni, no);
  return no;
}

// ---------------------------------------------------------------------------------------------------
// Euclidean clustering (euclidean_clustering.cpp:113-189)
// ---------------------------------------------------------------------------------------------------
std::vector<ClusterIndices> EuclideanClustering::extract() const {
  std::vector<ClusterIndices> result;
  if (!cloud_ || cloud_->n == 0) return result;
  const PointCloudSoA &c = *cloud_;
  const std::size_t n = c.n;
  float mn[3] = {c.x[0], c.y[0], c.z[0]}, mx[3] = {c.x[0], c.y[0], c.z[0]};
  for (std::size_t i = 0; i < n; ++i) {
    const float v[3] = {c.x[i], c.y[i], c.z[i]};
    for (int a = 0; a < 3; ++a) {
      if (v[a] < mn[a]) mn[a] = v[a];
      if (v[a] > mx[a]) mx[a] = v[a];
    }
  }
  float ext = 0.0f;
  for (int a = 0; a < 3; ++a)
    if (mx[a] - mn[a] > ext) ext = mx[a] - mn[a];
  // the grid edge only sets the query window; a zero tolerance still needs a positive edge
  const float cell = clusterTolerance_ > ext * 1e-6f ? clusterTolerance_ : (ext > 0.0f ? ext * 1e-6f : 1.0f);
  DevCloud d = upload_soa(c.x, c.y, c.z, n);
  rvp_hash *h = nullptr;
  check(rvp_hash_create(&h), "rvp_hash_create");
  struct HashGuard {
    rvp_hash *h;
    ~HashGuard() { rvp_hash_destroy(h); }
  } guard{h};
  check(rvp_hash_build(h, d.x.as<float>(), d.y.as<float>(), d.z.as<float>(), n, cell, nullptr), "rvp_hash_build");
  Dev idx(n * sizeof(int32_t)), off((n + 1) * sizeof(int64_t));
  std::size_t m = 0;
  check(rvp_euclidean_cluster(h, clusterTolerance_, minClusterSize_, maxClusterSize_, idx.as<int32_t>(),
                              off.as<int64_t>(), &m, nullptr, nullptr),
        "rvp_euclidean_cluster");
  std::vector<int64_t> ho(m + 1);
  download(ho.data(), off.p, (m + 1) * sizeof(int64_t));
  std::vector<int32_t> hi((std::size_t)ho[m]);
  if (!hi.empty()) download(hi.data(), idx.p, hi.size() * sizeof(int32_t));
  result.resize(m);
  for (std::size_t k = 0; k < m; ++k) result[k].indices.assign(hi.begin() + ho[k], hi.begin() + ho[k + 1]);
  return result;
}

// ---------------------------------------------------------------------------------------------------
// PCD I/O (simple_pcd_loader.h:143-328)
// ---------------------------------------------------------------------------------------------------
int loadPCD(const std::string &file_path, std::vector<PointXYZ> &points) {
  std::size_t n = 0;
  int mode = -1;
  if (rvp_pcd_info(file_path.c_str(), &n, &mode) != RVP_OK) return -1;
  std::vector<float> x(n ? n : 1), y(n ? n : 1), z(n ? n : 1);
  std::size_t m = 0;
  if (rvp_pcd_read(file_path.c_str(), x.data(), y.data(), z.data(), n, &m) != RVP_OK) return -1;
  points.resize(m);
  for (std::size_t i = 0; i < m; ++i) points[i] = {x[i], y[i], z[i]};
  return static_cast<int>(m);
}

void savePCD(const std::string &filename, const std::vector<PointXYZ> &points, bool binary) {
  std::vector<float> x(points.size()), y(points.size()), z(points.size());
  for (std::size_t i = 0; i < points.size(); ++i) {
    x[i] = points[i].x;
    y[i] = points[i].y;
    z[i] = points[i].z;
  }
  (void)rvp_pcd_write(filename.c_str(), x.data(), y.data(), z.data(), points.size(), binary ? 1 : 0);
}

}  // namespace rvpoint
