// walkthrough_input.cpp -- writes the cloud that the reference's pipeline walkthrough
// (tests/integration/test_pipeline_walkthrough.cpp:48-100) processes: the PCD file given as argv[1], loaded with
// loadPCD, plus its 1,000 synthetic ground points drawn the same way (std::mt19937(42), uniform x and y over the
// cloud's box widened by 0.02, z = min z - 0.005), so a test can run the oracle on exactly that input.
// Usage: walkthrough_input <in.pcd> <out.bin>; out.bin holds n (uint64) then n x {x, y, z} float32.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "rvv_pcl.h"

using namespace rvv_pcl;

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  std::vector<PointXYZ> pts;
  const int n = loadPCD(argv[1], pts);
  if (n <= 0) return 1;
  std::vector<float> x(n), y(n), z(n);
  for (int i = 0; i < n; ++i) {
    x[i] = pts[i].x;
    y[i] = pts[i].y;
    z[i] = pts[i].z;
  }
  const float min_z = *std::min_element(z.begin(), z.end());
  const float min_x = *std::min_element(x.begin(), x.end()), max_x = *std::max_element(x.begin(), x.end());
  const float min_y = *std::min_element(y.begin(), y.end()), max_y = *std::max_element(y.begin(), y.end());
  const float ground_z = min_z - 0.005f;
  std::mt19937 gen(42);
  std::uniform_real_distribution<float> dist_x(min_x - 0.02f, max_x + 0.02f);
  std::uniform_real_distribution<float> dist_y(min_y - 0.02f, max_y + 0.02f);
  for (int i = 0; i < 1000; ++i) {
    x.push_back(dist_x(gen));  // x before y, one draw each, as the walkthrough does
    y.push_back(dist_y(gen));
    z.push_back(ground_z);
  }
  FILE *f = std::fopen(argv[2], "wb");
  if (!f) return 1;
  const uint64_t m = x.size();
  std::fwrite(&m, sizeof(m), 1, f);
  for (uint64_t i = 0; i < m; ++i) {
    const float p[3] = {x[i], y[i], z[i]};
    std::fwrite(p, sizeof(p), 1, f);
  }
  std::fclose(f);
  return 0;
}
