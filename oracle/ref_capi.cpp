// ref_capi.cpp -- extern "C" wrappers around the reference's own scalar path, compiled from the reference
// sources where they lie (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle
// restatement and by bench.py's cpu_baseline leg.  Output goes to oracle/_ref/ only.
#include "core/point_types.h"
#include "core/rvv_common.h"
#include "search/spatial_hashing.h"
#include "search/octree.h"
#include "search/pointer_octree.h"
#include "search/radius_search.h"
#include "filters/voxel_grid.h"
#include "filters/statistical_outlier_removal.h"
#include "features/normal_estimation.h"
#include "io/simple_pcd_loader.h"
#include "segmentation/euclidean_clustering.h"
#include "segmentation/ransac_plane.h"

#include <cstdint>
#include <vector>

using namespace rvpoint;

static PointCloudSoA soa(const float *x, const float *y, const float *z, size_t n) {
  PointCloudSoA c;
  c.x = const_cast<float *>(x); c.y = const_cast<float *>(y); c.z = const_cast<float *>(z); c.n = n;
  return c;
}

extern "C" {

void ref_get_dist_sq(const float *x, const float *y, const float *z, float qx, float qy, float qz,
                     float *out, size_t n) {
  get_dist_sq_rvv(x, y, z, qx, qy, qz, out, n);
}

// get_inds_in_radius_rvv appending to vectors that already hold `have` entries (copied in first); returns the
// new length (rvv_common.cpp:49-68)
size_t ref_inds_in_radius(const float *x, const float *y, const float *z, const int32_t *subset, size_t m, float qx,
                          float qy, float qz, float r2, int32_t *out_idx, float *out_d2, size_t have) {
  std::vector<int> idx(out_idx, out_idx + have);
  std::vector<float> d2(out_d2, out_d2 + have);
  get_inds_in_radius_rvv(x, y, z, subset, m, qx, qy, qz, r2, idx, d2);
  for (size_t i = 0; i < idx.size(); ++i) {
    out_idx[i] = idx[i];
    out_d2[i] = d2[i];
  }
  return idx.size();
}

// SpatialHash::setInputCloud + build, then radiusSearch per query (spatial_hashing.cpp:76-223)
void ref_hash_radius(const float *x, const float *y, const float *z, size_t n, float cell,
                     const float *qx, const float *qy, const float *qz, size_t nq, float radius, int max_nn,
                     size_t stride, int32_t *idx_out, float *d2_out, int32_t *counts) {
  SpatialHash h;
  h.setInputCloud(soa(x, y, z, n), cell);
  h.build();
  std::vector<int> idx;
  std::vector<float> d;
  for (size_t q = 0; q < nq; ++q) {
    PointXYZ p{qx[q], qy[q], qz[q]};
    size_t m = h.radiusSearch(p, radius, idx, d, max_nn);
    counts[q] = (int32_t)m;
    for (size_t j = 0; j < m && j < stride; ++j) {
      if (idx_out) idx_out[q * stride + j] = idx[j];
      if (d2_out) d2_out[q * stride + j] = d[j];
    }
  }
}

size_t ref_radius_search_sc(const float *aos, size_t n, float qx, float qy, float qz, float r,
                            int32_t *idx, float *d, int max_nn) {
  return radius_search_sc(reinterpret_cast<const PointXYZ *>(aos), n, PointXYZ{qx, qy, qz}, r, idx, d, max_nn);
}

size_t ref_radius_search_rvv(const float *x, const float *y, const float *z, size_t n, float qx, float qy,
                             float qz, float r, int32_t *idx, float *d, int max_nn) {
  return radius_search_rvv(soa(x, y, z, n), PointXYZ{qx, qy, qz}, r, idx, d, max_nn);
}

size_t ref_sor_sc(const float *aos, size_t n, float *out, int k, float alpha) {
  return sor_sc(reinterpret_cast<const PointXYZ *>(aos), n, reinterpret_cast<PointXYZ *>(out), k, alpha);
}

size_t ref_sor_rvv(const float *x, const float *y, const float *z, size_t n, float *out, int k, float alpha) {
  return sor_rvv(soa(x, y, z, n), reinterpret_cast<PointXYZ *>(out), k, alpha);
}

size_t ref_sor_spatial_hash(const float *x, const float *y, const float *z, size_t n, float cell, float *out,
                            int k, float alpha, float search_radius) {
  SpatialHash h;
  PointCloudSoA c = soa(x, y, z, n);
  h.setInputCloud(c, cell);
  h.build();
  return sor_spatial_hash(c, h, reinterpret_cast<PointXYZ *>(out), k, alpha, search_radius);
}

size_t ref_sor_octree(const float *x, const float *y, const float *z, size_t n, float *out, int k, float alpha,
                      float search_radius, int pointer_tree) {
  PointCloudSoA c = soa(x, y, z, n);
  if (pointer_tree) {
    PointerOctree t;
    t.setInputCloud(c);
    t.build();
    return sor_pointer_octree(c, t, reinterpret_cast<PointXYZ *>(out), k, alpha, search_radius);
  }
  Octree t;
  t.setInputCloud(c);
  t.build();
  return sor_octree(c, t, reinterpret_cast<PointXYZ *>(out), k, alpha, search_radius);
}

// one query's neighbours from the reference's Octree (pointer_tree 0) or PointerOctree (1), traversal order
size_t ref_octree_radius(const float *x, const float *y, const float *z, size_t n, float qx, float qy, float qz,
                         float r, int pointer_tree, int *idx, float *d2, size_t cap) {
  PointCloudSoA c = soa(x, y, z, n);
  std::vector<int> vi;
  std::vector<float> vd;
  if (pointer_tree) {
    PointerOctree t;
    t.setInputCloud(c);
    t.build();
    t.radiusSearch(PointXYZ{qx, qy, qz}, r, vi, vd);
  } else {
    Octree t;
    t.setInputCloud(c);
    t.build();
    t.radiusSearch(PointXYZ{qx, qy, qz}, r, vi, vd);
  }
  for (size_t i = 0; i < vi.size() && i < cap; ++i) {
    idx[i] = vi[i];
    d2[i] = vd[i];
  }
  return vi.size();
}

void ref_normal_octree(const float *x, const float *y, const float *z, size_t n, float *nx, float *ny, float *nz,
                       int k, float radius, float vpx, float vpy, float vpz, int iters) {
  PointCloudSoA c = soa(x, y, z, n);
  Octree t;
  t.setInputCloud(c);
  t.build();
  normal_estimation_rvv(c, t, nx, ny, nz, k, radius, vpx, vpy, vpz, iters);
}

size_t ref_voxel_sc(const float *aos, size_t n, float *out, float leaf) {
  return voxel_grid_downsamp_sc(reinterpret_cast<const PointXYZ *>(aos), n, reinterpret_cast<PointXYZ *>(out), leaf);
}

size_t ref_voxel_v2(const float *x, const float *y, const float *z, size_t n, float *out, float leaf) {
  return voxel_grid_downsamp_rvv_v2(soa(x, y, z, n), reinterpret_cast<PointXYZ *>(out), leaf);
}

void ref_normal_sc(const float *aos, size_t n, float *nx, float *ny, float *nz, int k, float radius,
                   float vpx, float vpy, float vpz, int iters) {
  normal_estimation_sc(reinterpret_cast<const PointXYZ *>(aos), n, nx, ny, nz, k, radius, vpx, vpy, vpz, iters);
}

void ref_normal_hash(const float *x, const float *y, const float *z, size_t n, float cell, float *nx, float *ny,
                     float *nz, int k, float radius, float vpx, float vpy, float vpz, int iters) {
  SpatialHash h;
  PointCloudSoA c = soa(x, y, z, n);
  h.setInputCloud(c, cell);
  h.build();
  normal_estimation_rvv(c, h, nx, ny, nz, k, radius, vpx, vpy, vpz, iters);
}

int ref_ransac_sc(const float *aos, size_t n, float thr, int max_iters, float *model, float collinear, float prob) {
  return ransac_plane_sc(reinterpret_cast<const PointXYZ *>(aos), n, thr, max_iters, model, collinear, prob);
}

int ref_ransac_rvv(const float *x, const float *y, const float *z, size_t n, float thr, int max_iters,
                   float *model, float collinear, float prob) {
  return ransac_plane_rvv(soa(x, y, z, n), thr, max_iters, model, collinear, prob);
}

void ref_extract(const float *x, const float *y, const float *z, size_t n, const float *model, float thr,
                 float *inl, float *outl, size_t *ni, size_t *no) {
  extract_plane_inliers_outliers_rvv(soa(x, y, z, n), model, thr, reinterpret_cast<PointXYZ *>(inl),
                                     reinterpret_cast<PointXYZ *>(outl), *ni, *no);
}

size_t ref_euclidean_cluster(const float *x, const float *y, const float *z, size_t n, float tol, int min_size,
                             int max_size, int32_t *indices, int64_t *offsets) {
  rvpoint::PointCloudSoA c{const_cast<float *>(x), const_cast<float *>(y), const_cast<float *>(z), n};
  rvpoint::EuclideanClustering ec;
  ec.setInputCloud(c);
  ec.setClusterTolerance(tol);
  ec.setMinClusterSize(min_size);
  ec.setMaxClusterSize(max_size);
  const auto clusters = ec.extract();
  size_t out = 0;
  offsets[0] = 0;
  for (size_t i = 0; i < clusters.size(); ++i) {
    for (int v : clusters[i].indices) indices[out++] = v;
    offsets[i + 1] = (int64_t)out;
  }
  return clusters.size();
}

int ref_load_pcd(const char *path, float *aos, size_t cap) {
  std::vector<rvpoint::PointXYZ> pts;
  const int n = rvpoint::loadPCD(path, pts);
  if (n < 0) return n;
  for (size_t i = 0; i < pts.size() && i < cap; ++i) {
    aos[3 * i] = pts[i].x;
    aos[3 * i + 1] = pts[i].y;
    aos[3 * i + 2] = pts[i].z;
  }
  return n;
}

void ref_save_pcd(const char *path, const float *aos, size_t n, int binary) {
  std::vector<rvpoint::PointXYZ> pts(n);
  for (size_t i = 0; i < n; ++i) pts[i] = {aos[3 * i], aos[3 * i + 1], aos[3 * i + 2]};
  rvpoint::savePCD(path, pts, binary != 0);
}

}  // extern "C"
