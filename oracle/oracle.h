/*
 * oracle.h -- CPU restatement of RVPoint's scalar neighbour-search path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (rvpoint_amd/, include/)
 * may link, load or call this library; it exists so that tests/, the smoke
 * check and the CPU-baseline leg of bench.py can check the HIP path against
 * an independent, plain-C statement of the reference's algorithm.
 *
 * Every function cites the reference file:line whose operation order it
 * follows.  Build: gcc -O2 -ffp-contract=off (no -ffast-math), so every fp32
 * operation rounds once, in the order written.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_bbox / orc_cell_coords      closed-form cells at hand-picked edges; oracle/_ref
 *   orc_hash_radius                 brute force (orc_radius_all) on random clouds; oracle/_ref
 *   orc_radius_bruteforce           oracle/_ref radius_search_sc; hand-counted tiny cases
 *   orc_knn                         full sort of all distances (tiny inputs)
 *   orc_sor_sc / orc_sor_radius     oracle/_ref sor_sc / sor_spatial_hash; KAT counts
 *   orc_voxel_sc                    oracle/_ref voxel_grid_downsamp_sc; KAT counts 18,542 / 52,800
 *   orc_normals_*_f64               analytic plane/sphere normals (closed form)
 *   orc_eigen_e1                    oracle/_ref normal_estimation_sc on the same neighbourhoods
 *   orc_rand_*                      glibc rand() of the host (ctypes)
 *   orc_ransac_plane                oracle/_ref ransac_plane_sc (count + model, bit-exact)
 *   orc_plane_extract               oracle/_ref extract_plane_inliers_outliers_rvv
 *   orc_euclidean_cluster           oracle/_ref EuclideanClustering::extract; union-find components (numpy)
 */
#ifndef RVP_ORACLE_H
#define RVP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* H1: bbox with the 0.01*cell margin (spatial_hashing.cpp:76-146 scalar branch). out = {minx,miny,minz,maxx,maxy,maxz} */
/* D1 / D2 (rvv_common.cpp:36-43, :55-68); D2 appends after `have` entries and returns the new length */
void orc_dist_sq(const float *x, const float *y, const float *z, size_t n, float qx, float qy, float qz, float *out);
size_t orc_inds_in_radius(const float *x, const float *y, const float *z, const int32_t *subset, size_t m,
                          float qx, float qy, float qz, float r2, int32_t *out_idx, float *out_d2, size_t have);
void orc_bbox(const float *x, const float *y, const float *z, size_t n, float cell, float out[6]);

/* H2: per-point cell coordinates (spatial_hashing.h:47-52) and the int64 hash key (spatial_hashing.h:43-45). */
void orc_cell_coords(const float *x, const float *y, const float *z, size_t n, float cell,
                     int32_t *ixyz /* n*3 */, int64_t *keys /* n, may be NULL */);

/* H3: SpatialHash::setInputCloud+build+radiusSearch for nq queries (spatial_hashing.cpp:76-223).
 * Results for query i go to idx_out[i*row_stride ..], d2_out likewise, count to counts[i].
 * The full (untruncated) neighbour count is written to total[i] when total != NULL.
 * max_nn <= 0 means unlimited; row_stride must be >= the longest list then. */
void orc_hash_radius(const float *x, const float *y, const float *z, size_t n, float cell,
                     const float *qx, const float *qy, const float *qz, size_t nq,
                     float radius, int max_nn, size_t row_stride,
                     int32_t *idx_out, float *d2_out, int32_t *counts, int32_t *total);

/* Brute-force fixed-radius set sorted by (d2, idx) -- the set H3 must equal. Same output layout as orc_hash_radius. */
void orc_radius_all(const float *x, const float *y, const float *z, size_t n,
                    const float *qx, const float *qy, const float *qz, size_t nq,
                    float radius, int max_nn, size_t row_stride,
                    int32_t *idx_out, float *d2_out, int32_t *counts, int32_t *total);

/* H4: radius_search_sc (radius_search.cpp:12-33): first max_nn hits in index order. Returns min(count, max_nn). */
size_t orc_radius_bruteforce(const float *x, const float *y, const float *z, size_t n,
                             float qx, float qy, float qz, float radius,
                             int32_t *indices, float *dists, int max_nn);

/* k nearest by (d2, idx) over all n points, query order q - p (statistical_outlier_removal.cpp:27-34). */
void orc_knn(const float *x, const float *y, const float *z, size_t n,
             const float *qx, const float *qy, const float *qz, size_t nq,
             int k, int32_t *idx_out /* nq*k */, float *d2_out /* nq*k */);

/* S1: sor_sc (statistical_outlier_removal.cpp:19-63). aos = n*3 floats. Writes kept points (AoS) to out,
 * per-point mean kNN distance to means (may be NULL). Returns the kept count. */
size_t orc_sor_sc(const float *aos, size_t n, int k, float alpha, float *out, float *means);

/* S2: sor_spatial_hash + filter_by_mean_dists (statistical_outlier_removal.cpp:246-270, 193-218),
 * hash built on the same cloud with cell size `cell`. */
/* orc_sor_sc with the per-point means on `threads` OpenMP threads (global sums sequential; same bits) */
size_t orc_sor_sc_mt(const float *aos, size_t n, int k, float alpha, float *out, float *means, int threads);
size_t orc_sor_radius(const float *x, const float *y, const float *z, size_t n, float cell,
                      int k, float alpha, float search_radius, float *out, float *means);

/* Global-statistics tail shared by S1/S2: sequential fp32 sums in index order. thr_out = gmean + alpha*std. */
size_t orc_sor_filter(const float *x, const float *y, const float *z, size_t n,
                      const float *means, float alpha, float *out, uint8_t *keep, float stats[3]);

/* Sequential fp32 sums of sor_sc / filter_by_mean_dists (statistical_outlier_removal.cpp:43-52, 193-205):
 * sum of v (squared_dev == 0) or of fl(fl(v - mean)^2), left to right from +0. */
float orc_seq_sum(const float *v, size_t n, int squared_dev, float mean);

/* V1: voxel_grid_downsamp_sc (voxel_grid_downsamp.cpp:183-213). aos = n*3. Output AoS in (vx,vy,vz) order. */
size_t orc_voxel_sc(const float *aos, size_t n, float leaf, float *out);

/* N1 semantics, FP64 arithmetic: neighbourhood = k+1 nearest incl. self by (d2, idx); fp64 centroid,
 * covariance, Jacobi eigen solve to convergence; flip toward the viewpoint (normal_estimation.cpp:72-112, 55-67). */
void orc_normals_knn_f64(const float *x, const float *y, const float *z, size_t n, int k,
                         const float vp[3], double *nx, double *ny, double *nz);

/* N2 semantics, FP64 arithmetic: neighbourhood = H3 list (cap max_nn, 0 = unlimited) with < 3 -> zero normal
 * (normal_estimation.cpp:189-214). */
void orc_normals_radius_f64(const float *x, const float *y, const float *z, size_t n, float cell,
                            float radius, int max_nn, const float vp[3],
                            double *nx, double *ny, double *nz);

/* FP64 covariance + eigen solve for one neighbourhood given as coordinates; returns smallest-eigenvalue
 * eigenvector (unit) and the three eigenvalues ascending. */
void orc_normal_from_points_f64(const double *px, const double *py, const double *pz, size_t m,
                                double n_out[3], double evals[3]);

/* E1: simple_eigen3x3_smallest restated (normal_estimation.cpp:16-52).  Reference behaviour only -- it is
 * NOT an accuracy oracle (wrong rotation sign, unrotated off-diagonals). */
void orc_eigen_e1(const float cov[9], int iters, float n_out[3]);

/* F1: flipNormalTowardsViewpoint (normal_estimation.cpp:55-67), fp32. */
/* normal_estimation_rvv with a SpatialHash (reference semantics: uncapped radius list, E1, flip), threaded */
void orc_normals_hash_e1_mt(const float *x, const float *y, const float *z, size_t n, float cell, float radius,
                            const float vp[3], int eigen_iters, float *nx, float *ny, float *nz, int threads);
void orc_flip(float px, float py, float pz, float vx, float vy, float vz, float n[3]);

/* glibc TYPE_3 random_r restatement (srand/rand). */
typedef struct { int32_t r[34]; int pos; } orc_rand_state;
void orc_srand(orc_rand_state *s, unsigned seed);
int32_t orc_rand(orc_rand_state *s);

/* P1: compute_plane_coefficients (ransac_plane.cpp:18-50). Returns 1 and writes model, or 0 when collinear. */
int orc_plane_model(const float p1[3], const float p2[3], const float p3[3], float collinear, float model[4]);

/* Inlier count of one model, fabsf(((a*x+b*y)+c*z)+d) <= thr (ransac_plane.cpp:81-89). */
int orc_plane_count(const float *x, const float *y, const float *z, size_t n, const float model[4], float thr);

/* R1: ransac_plane_sc (ransac_plane.cpp:55-111) with the restated glibc generator. Returns best inliers.
 * iters_run (may be NULL) receives the number of loop iterations executed. */
int orc_ransac_plane(const float *x, const float *y, const float *z, size_t n, float thr, int max_iters,
                     float collinear, float probability, float model[4], int *iters_run);

/* X1: extract_plane_inliers_outliers scalar branch (ransac_plane.cpp:374-382). AoS outputs. */
void orc_plane_extract(const float *x, const float *y, const float *z, size_t n, const float model[4],
                       float thr, float *inl, size_t *n_in, float *outl, size_t *n_out);

/* Euclidean clustering, scalar branch without a search structure (euclidean_clustering.cpp:49-111,
 * 113-189): breadth-first search from every unvisited seed in index order; the neighbours of a point q are
 * the unvisited i with ((dx*dx + dy*dy) + dz*dz) <= fl(tol*tol), dx = x_i - q_x, each op rounded once;
 * clusters of min_size..max_size points are kept, their indices sorted.  CSR output: indices[<= n],
 * offsets[clusters + 1]; returns the number of clusters. */
size_t orc_euclidean_cluster(const float *x, const float *y, const float *z, size_t n, float tol, int min_size,
                             int max_size, int32_t *indices, int64_t *offsets);

#ifdef __cplusplus
}
#endif
#endif
