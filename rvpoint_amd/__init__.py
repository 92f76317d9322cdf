"""rvpoint_amd -- MI355X (gfx950) neighbour-search path of RVPoint: hash-grid build, radius / k-NN
queries, SOR, normals, voxel grid and RANSAC plane counting, behind the reference's API names."""
from ._lib import RvpError, gpu_available, lib  # noqa: F401
from .api import (  # noqa: F401
    Batch,
    batch_normals_knn,
    batch_normals_radius,
    batch_radius_search,
    batch_ransac_plane,
    batch_sor_exact,
    batch_voxel,
    PointCloudSoA,
    SpatialHash,
    extract_plane_inliers_outliers_rvv,
    extract_plane_inliers_rvv,
    extract_plane_outliers_rvv,
    normal_estimation_rvv,
    normal_estimation_sc,
    normals_fallback_count,
    radius_search_batch_bruteforce,
    radius_search_rvv,
    radius_search_sc,
    ransac_plane_rvv,
    ransac_plane_sc,
    sor_rvv,
    sor_sc,
    sor_spatial_hash,
    voxel_grid_downsamp_rvv,
    voxel_grid_downsamp_rvv_v2,
    voxel_grid_downsamp_sc,
)
from .device import DeviceArray, DeviceCloud, Stream, synchronize  # noqa: F401
