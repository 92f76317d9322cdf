// rvv_pcl.h -- umbrella header of the drop-in API: the reference's rvpoint:: signatures for the
// neighbour-search path, also visible as namespace rvv_pcl (the name the reference's tests use).
#pragma once

#include "core/point_types.h"
#include "core/rvv_common.h"
#include "features/normal_estimation.h"
#include "io/simple_pcd_loader.h"
#include "filters/statistical_outlier_removal.h"
#include "filters/voxel_grid.h"
#include "search/octree.h"
#include "search/pointer_octree.h"
#include "search/radius_search.h"
#include "search/spatial_hashing.h"
#include "segmentation/euclidean_clustering.h"
#include "segmentation/ransac_plane.h"

namespace rvv_pcl {
using namespace rvpoint;
}
