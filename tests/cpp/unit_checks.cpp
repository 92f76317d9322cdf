// unit_checks.cpp -- the checks of the reference's own unit tests (tests/unit/test_radius.cpp,
// test_sor.cpp, test_normal.cpp, test_voxel_grid.cpp, test_ransac.cpp), run against the drop-in API
// (include/rvv_pcl.h) with the same std::mt19937(42) inputs, plus the D1 / D2 helpers of rvv_common.h.
// Exit status 0 = all pass.  The normals check uses the SpatialHash overload of normal_estimation_rvv (the
// reference's test_normal.cpp itself, with its Octree, runs unchanged as its own program).
// With "--dump <file>", the D1 / D2 inputs and outputs are also written there (raw little-endian arrays) so that
// a test can compare the C++ layer with the oracle.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <iostream>
#include <random>
#include <set>
#include <vector>

#include "rvv_pcl.h"

using namespace rvv_pcl;

static int fails = 0;
#define EXPECT(cond, what)                                      \
  do {                                                          \
    if (!(cond)) {                                              \
      std::cerr << "[FAIL] " << what << std::endl;              \
      ++fails;                                                  \
    } else {                                                    \
      std::cout << "[PASS] " << what << std::endl;              \
    }                                                           \
  } while (0)

static bool close3(const PointXYZ &a, const PointXYZ &b, float eps = 1e-4f) {
  return std::abs(a.x - b.x) < eps && std::abs(a.y - b.y) < eps && std::abs(a.z - b.z) < eps;
}

static void radius_check() {  // tests/unit/test_radius.cpp:11-60
  const size_t N = 1000;
  const int MAX_NN = 100;
  std::vector<float> x(N), y(N), z(N);
  std::vector<PointXYZ> aos(N);
  std::mt19937 gen(42);
  std::normal_distribution<float> dist(0.0f, 2.0f);
  for (size_t i = 0; i < N; ++i) {
    x[i] = dist(gen); y[i] = dist(gen); z[i] = dist(gen);
    aos[i] = {x[i], y[i], z[i]};
  }
  PointCloudSoA soa = {x.data(), y.data(), z.data(), N};
  std::vector<int> i1(MAX_NN), i2(MAX_NN);
  std::vector<float> d1(MAX_NN), d2(MAX_NN);
  size_t c1 = radius_search_sc(aos.data(), N, {0, 0, 0}, 2.0f, i1.data(), d1.data(), MAX_NN);
  size_t c2 = radius_search_rvv(soa, {0, 0, 0}, 2.0f, i2.data(), d2.data(), MAX_NN);
  EXPECT(c1 == c2 && c1 > 0, "radius: sc and rvv counts equal");
  EXPECT(std::set<int>(i1.begin(), i1.begin() + c1) == std::set<int>(i2.begin(), i2.begin() + c2),
         "radius: index sets equal");
  // SpatialHash agrees with brute force on the same query
  SpatialHash h;
  h.setInputCloud(soa, 0.5f);
  h.build();
  std::vector<int> hi;
  std::vector<float> hd;
  h.radiusSearch({0, 0, 0}, 2.0f, hi, hd);
  std::vector<int> all(N);
  std::vector<float> alld(N);
  size_t ca = radius_search_sc(aos.data(), N, {0, 0, 0}, 2.0f, all.data(), alld.data(), (int)N);
  EXPECT(std::set<int>(hi.begin(), hi.end()) == std::set<int>(all.begin(), all.begin() + ca),
         "radius: SpatialHash set equals brute force");
  EXPECT(std::is_sorted(hd.begin(), hd.end()), "radius: SpatialHash list sorted by distance");
}

static void sor_check() {  // tests/unit/test_sor.cpp:10-60
  const size_t N = 1000;
  std::vector<float> x(N), y(N), z(N);
  std::vector<PointXYZ> aos(N);
  std::mt19937 gen(42);
  std::normal_distribution<float> cluster(0.0f, 1.0f);
  std::uniform_real_distribution<float> outl(-20.0f, 20.0f);
  for (size_t i = 0; i < N; ++i) {
    if (i < N * 0.9) { x[i] = cluster(gen); y[i] = cluster(gen); z[i] = cluster(gen); }
    else { x[i] = outl(gen); y[i] = outl(gen); z[i] = outl(gen); }
    aos[i] = {x[i], y[i], z[i]};
  }
  PointCloudSoA soa = {x.data(), y.data(), z.data(), N};
  std::vector<PointXYZ> o1(N), o2(N);
  size_t c1 = sor_sc(aos.data(), N, o1.data(), 10, 1.0f);
  size_t c2 = sor_rvv(soa, o2.data(), 10, 1.0f);
  bool same = c1 == c2;
  for (size_t i = 0; same && i < c1; ++i) same = close3(o1[i], o2[i]);
  EXPECT(same && c1 > 0 && c1 < N, "sor: sc and rvv agree");
  SpatialHash h;
  h.setInputCloud(soa, 0.5f);
  h.build();
  std::vector<PointXYZ> o3(N);
  size_t c3 = sor_spatial_hash(soa, h, o3.data(), 10, 1.0f, 0.5f);
  EXPECT(c3 > 0 && c3 <= N, "sor: spatial-hash variant runs");
}

static void normal_check() {  // tests/unit/test_normal.cpp:10-60 (SpatialHash in place of Octree)
  const size_t N = 1000;
  std::vector<float> x(N), y(N), z(N);
  std::vector<PointXYZ> aos(N);
  std::mt19937 gen(42);
  std::uniform_real_distribution<float> surf(-10.0f, 10.0f);
  std::normal_distribution<float> noise(0.0f, 0.01f);
  for (size_t i = 0; i < N; ++i) {
    x[i] = surf(gen); y[i] = surf(gen); z[i] = noise(gen);
    aos[i] = {x[i], y[i], z[i]};
  }
  PointCloudSoA soa = {x.data(), y.data(), z.data(), N};
  std::vector<float> a(N), b(N), c(N), p(N), q(N), r(N);
  normal_estimation_sc(aos.data(), N, a.data(), b.data(), c.data(), 10, 2.0f);
  SpatialHash h;
  h.setInputCloud(soa, 2.0f);
  h.build();
  normal_estimation_rvv(soa, h, p.data(), q.data(), r.data(), 10, 2.0f);
  size_t pass = 0;
  for (size_t i = 0; i < N; ++i) {
    float dot = a[i] * p[i] + b[i] * q[i] + c[i] * r[i];
    if (std::abs(c[i]) > 0.9f && std::abs(r[i]) > 0.9f && std::abs(dot) > 0.9f) ++pass;
  }
  EXPECT(pass >= N * 0.95, "normal: >= 95% of normals agree and are vertical");
}

static void voxel_check() {  // tests/unit/test_voxel_grid.cpp:10-80
  const size_t N = 1000;
  std::vector<float> x(N), y(N), z(N);
  std::vector<PointXYZ> aos(N);
  std::mt19937 gen(42);
  std::uniform_real_distribution<float> dist(0.0f, 10.0f);
  for (size_t i = 0; i < N; ++i) {
    x[i] = dist(gen); y[i] = dist(gen); z[i] = dist(gen);
    aos[i] = {x[i], y[i], z[i]};
  }
  PointCloudSoA soa = {x.data(), y.data(), z.data(), N};
  std::vector<PointXYZ> o1(N), o2(N), o3(N);
  size_t c1 = voxel_grid_downsamp_sc(aos.data(), N, o1.data(), 0.5f);
  size_t c2 = voxel_grid_downsamp_rvv(soa, o2.data(), 0.5f);
  size_t c3 = voxel_grid_downsamp_rvv_v2(soa, o3.data(), 0.5f);
  EXPECT(c1 == c2 && c1 == c3 && c1 > 0, "voxel: counts equal");
  auto lt = [](const PointXYZ &u, const PointXYZ &v) {
    if (u.x != v.x) return u.x < v.x;
    if (u.y != v.y) return u.y < v.y;
    return u.z < v.z;
  };
  std::sort(o1.begin(), o1.begin() + c1, lt);
  std::sort(o2.begin(), o2.begin() + c2, lt);
  std::sort(o3.begin(), o3.begin() + c3, lt);
  bool same = true;
  for (size_t i = 0; i < c1; ++i) same = same && close3(o1[i], o2[i]) && close3(o1[i], o3[i]);
  EXPECT(same, "voxel: centroids agree");
}

static void ransac_check() {  // tests/unit/test_ransac.cpp:11-64
  const size_t NI = 500, NO = 500, N = NI + NO;
  std::vector<float> x(N), y(N), z(N);
  std::vector<PointXYZ> aos(N);
  std::mt19937 gen(42);
  std::uniform_real_distribution<float> rxy(-10, 10);
  std::normal_distribution<float> rz(10.0f, 0.01f);
  std::uniform_real_distribution<float> ro(-20, 20);
  for (size_t i = 0; i < N; ++i) {
    if (i < NI) { x[i] = rxy(gen); y[i] = rxy(gen); z[i] = rz(gen); }
    else { x[i] = ro(gen); y[i] = ro(gen); z[i] = ro(gen); }
    aos[i] = {x[i], y[i], z[i]};
  }
  PointCloudSoA soa = {x.data(), y.data(), z.data(), N};
  float m1[4], m2[4];
  int c1 = ransac_plane_sc(aos.data(), N, 0.1f, 1000, m1);
  int r_after = std::rand();
  int c2 = ransac_plane_rvv(soa, 0.1f, 1000, m2);
  EXPECT(c1 >= 450 && c2 >= 450, "ransac: >= 450 inliers");
  EXPECT(std::abs(m1[2]) > 0.9f && std::abs(m1[3]) >= 9.0f && std::abs(m2[2]) > 0.9f && std::abs(m2[3]) >= 9.0f,
         "ransac: model is the z = 10 plane");
  EXPECT(c1 == c2 && m1[0] == m2[0] && m1[3] == m2[3], "ransac: sc and rvv identical");
  // the C library generator is left where the reference leaves it: replay the same call and compare
  (void)ransac_plane_sc(aos.data(), N, 0.1f, 1000, m1);
  EXPECT(std::rand() == r_after, "ransac: rand() state after the call matches a deterministic replay");
  std::vector<PointXYZ> in(N), out(N);
  size_t ni = 0, no = 0;
  extract_plane_inliers_outliers_rvv(soa, m1, 0.1f, in.data(), out.data(), ni, no);
  EXPECT(ni + no == N && ni >= 450, "extract: inliers + outliers partition the cloud");
}

// D1 / D2 (rvv_common.cpp:36-68): get_inds_in_radius_rvv appends after what the vectors hold, in subset order,
// and its d2 values are get_dist_sq_rvv's.
static void distance_check(const char *dump) {
  const size_t N = 3000, M = 2000;
  std::vector<float> x(N), y(N), z(N);
  std::mt19937 gen(42);
  std::normal_distribution<float> dist(0.0f, 1.0f);
  for (size_t i = 0; i < N; ++i) { x[i] = dist(gen); y[i] = dist(gen); z[i] = dist(gen); }
  std::vector<int> subset(M);
  for (size_t i = 0; i < M; ++i) subset[i] = (int)((i * 7919u + 13u) % N);  // a permuted subset
  const float qx = 0.1f, qy = -0.2f, qz = 0.3f, r2 = 0.8f;
  std::vector<float> d(N);
  get_dist_sq_rvv(x.data(), y.data(), z.data(), qx, qy, qz, d.data(), N);
  std::vector<int> idx = {-1, -2, -3};
  std::vector<float> dd = {9.0f, 8.0f, 7.0f};
  get_inds_in_radius_rvv(x.data(), y.data(), z.data(), subset.data(), M, qx, qy, qz, r2, idx, dd);
  bool prefix = idx.size() == dd.size() && idx.size() >= 3 && idx[0] == -1 && idx[1] == -2 && idx[2] == -3 &&
                dd[0] == 9.0f && dd[1] == 8.0f && dd[2] == 7.0f;
  EXPECT(prefix, "distance: get_inds_in_radius_rvv appends without clearing");
  size_t j = 3;
  bool order = true;
  for (size_t i = 0; i < M; ++i)
    if (d[subset[i]] <= r2) {
      order = order && j < idx.size() && idx[j] == subset[i] && dd[j] == d[subset[i]];
      ++j;
    }
  EXPECT(order && j == idx.size(), "distance: hits in subset order with get_dist_sq_rvv's d2");
  if (dump) {
    FILE *f = std::fopen(dump, "wb");
    if (!f) { ++fails; return; }
    const uint64_t hdr[3] = {N, M, idx.size()};
    const float q[4] = {qx, qy, qz, r2};
    std::fwrite(hdr, sizeof(hdr), 1, f);
    std::fwrite(q, sizeof(q), 1, f);
    std::fwrite(x.data(), 4, N, f); std::fwrite(y.data(), 4, N, f); std::fwrite(z.data(), 4, N, f);
    std::fwrite(subset.data(), 4, M, f);
    std::fwrite(d.data(), 4, N, f);
    std::fwrite(idx.data(), 4, idx.size(), f);
    std::fwrite(dd.data(), 4, dd.size(), f);
    std::fclose(f);
  }
}

int main(int argc, char **argv) {
  const char *dump = (argc == 3 && std::strcmp(argv[1], "--dump") == 0) ? argv[2] : nullptr;
  distance_check(dump);
  radius_check();
  sor_check();
  normal_check();
  voxel_check();
  ransac_check();
  std::cout << (fails ? "[FAIL] " : "[PASS] ") << fails << " failure(s)" << std::endl;
  return fails ? 1 : 0;
}
