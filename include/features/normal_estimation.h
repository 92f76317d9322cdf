// normal_estimation.h -- normal estimation (src/features/normal_estimation.h:15-49), on the GPU.
// The eigen solve converges (closed form with an FP64 fallback); eigen_iters is accepted and unused.
#pragma once

#include <cstddef>

#include "core/point_types.h"

namespace rvpoint {

class SpatialHash;
class Octree;
class PointerOctree;

// k+1 nearest (incl. the point itself) per point; normals flipped toward (vp_x, vp_y, vp_z).
void normal_estimation_sc(const PointXYZ *in, std::size_t n, float *nx, float *ny, float *nz, int k, float radius,
                          float vp_x = 0, float vp_y = 0, float vp_z = 0, int eigen_iters = 4);

// All neighbours within radius from the hash; fewer than 3 gives a zero normal.
void normal_estimation_rvv(const PointCloudSoA &in, const SpatialHash &hash, float *nx, float *ny, float *nz,
                           int k, float radius, float vp_x = 0, float vp_y = 0, float vp_z = 0,
                           int eigen_iters = 4);

// The same over the tree facades (octree.h, pointer_octree.h), and self-contained (a grid of edge `radius`).
void normal_estimation_rvv(const PointCloudSoA &in, const Octree &octree, float *nx, float *ny, float *nz, int k,
                           float radius, float vp_x = 0, float vp_y = 0, float vp_z = 0, int eigen_iters = 4);
void normal_estimation_rvv(const PointCloudSoA &in, const PointerOctree &octree, float *nx, float *ny, float *nz,
                           int k, float radius, float vp_x = 0, float vp_y = 0, float vp_z = 0,
                           int eigen_iters = 4);
void normal_estimation_rvv(const PointCloudSoA &in, float *nx, float *ny, float *nz, int k, float radius,
                           float vp_x = 0, float vp_y = 0, float vp_z = 0, int eigen_iters = 4);

// Extension: the same with the (d2, idx)-sorted neighbour list capped at max_nn.
void normal_estimation_rvv_capped(const PointCloudSoA &in, const SpatialHash &hash, float *nx, float *ny,
                                  float *nz, float radius, int max_nn, float vp_x = 0, float vp_y = 0,
                                  float vp_z = 0);

}  // namespace rvpoint
