/*
 * rvp.h -- C ABI of the MI355X (gfx950) neighbour-search path: librvp_hip.so.
 *
 * Every array argument is a DEVICE pointer (HBM) unless its comment says "host"; scalar results that
 * the caller reads (counts, models) are host pointers and the call returns after they are valid.
 * Calls are ordered on the given HIP stream (NULL = the null stream).  Return value: RVP_OK (0) or a
 * negative RVP_E_* status; rvp_last_error() gives a message.  Nothing throws across this boundary.
 *
 * Points are SoA float x[n], y[n], z[n] (the reference's PointCloudSoA, src/core/point_types.h:24-29);
 * AoS outputs are float[3*n] laid out as the reference's PointXYZ (src/core/point_types.h:12-14).
 *
 * Each entry point names the reference function (file:line) whose results it reproduces.
 */
#ifndef RVP_H
#define RVP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  RVP_OK = 0,
  RVP_E_ARG = -1,        /* bad argument (NULL pointer, negative size, k out of range) */
  RVP_E_OOM = -2,        /* device allocation failed */
  RVP_E_DEVICE = -3,     /* HIP runtime / kernel error */
  RVP_E_STATE = -4,      /* handle used before build */
  RVP_E_UNSUPPORTED = -5 /* a size beyond what the kernels support */
};

typedef void *rvp_stream_t; /* hipStream_t */
typedef struct rvp_hash rvp_hash;

/* ---- runtime / memory helpers (so hosts need no HIP headers) ------------------------------------- */
const char *rvp_last_error(void);
const char *rvp_version(void);
int rvp_device_count(int *count);
int rvp_set_device(int device);
int rvp_malloc(void **ptr, size_t bytes);
int rvp_free(void *ptr);
int rvp_memcpy_h2d(void *dst, const void *src, size_t bytes, rvp_stream_t s);
int rvp_memcpy_d2h(void *dst, const void *src, size_t bytes, rvp_stream_t s);
int rvp_memcpy_d2d(void *dst, const void *src, size_t bytes, rvp_stream_t s);
int rvp_memset(void *dst, int value, size_t bytes, rvp_stream_t s);
int rvp_stream_create(rvp_stream_t *s);
int rvp_stream_destroy(rvp_stream_t s);
/* Temporaries come from a cache keyed by (device, stream) and stay cached between calls; trim() waits for
 * the device and returns every cached block of the current device to the driver (destroying a stream
 * returns that stream's blocks). */
int rvp_cache_trim(void);
int rvp_stream_sync(rvp_stream_t s);
int rvp_device_sync(void);
/* event timing on a stream: start/stop record, elapsed in ms */
int rvp_event_create(void **ev);
int rvp_event_destroy(void *ev);
int rvp_event_record(void *ev, rvp_stream_t s);
int rvp_event_elapsed(void *start, void *stop, float *ms);
/* later work on stream s waits for the last record of event ev (no counterpart in the reference, which has no
 * device): lets one stream consume what another produced without a host wait */
int rvp_stream_wait_event(rvp_stream_t s, void *ev);

/* AoS float[3n] -> SoA x/y/z (the copy the reference's drivers make between stages,
 * src/tools/pipeline_export.cpp:336-342) */
int rvp_aos_to_soa(const float *aos, size_t n, float *x, float *y, float *z, rvp_stream_t s);
/* The same for min(*n_dev, n_bound) points, the count read on the device (n_dev: a device pointer, e.g.
 * the last entry of a batched stage's output offsets), so a pipeline converts what a stage kept without a
 * read-back (no counterpart in the reference, whose drivers know their counts on the host). */
int rvp_aos_to_soa_bounded(const float *aos, size_t n_bound, const uint32_t *n_dev, float *x, float *y, float *z,
                           rvp_stream_t s);
/* Graphs (no counterpart in the reference, which has no device): capture the calls made on a created stream
 * between begin and end into one graph, and replay it on that stream.  The batched stages (rvp_batch_*, the
 * frame pipeline's voxel, SOR and normals; not RANSAC, whose adaptive stop is replayed on the host) record only
 * kernels: their look-back epochs live on the device and their temporaries come from an arena the graph keeps,
 * so every replay recomputes from the inputs' current contents.  Run the calls once eagerly on the same stream
 * before capturing (first-use allocations happen then).  Replays must not overlap with eager calls on other
 * streams that write the same buffers; destroy() waits for the stream. */
typedef struct rvp_graph rvp_graph;
int rvp_graph_capture_begin(rvp_stream_t s);
int rvp_graph_capture_end(rvp_stream_t s, rvp_graph **graph);
int rvp_graph_launch(rvp_graph *graph, rvp_stream_t s);
int rvp_graph_destroy(rvp_graph *graph);
/* Per-kernel timing: while enabled, every kernel launch of the library is bracketed by a HIP event pair on
 * the stream it runs on.  collect() synchronises, writes "kernel launches total_ms" lines into buf
 * (NUL-terminated, *needed = size required) and clears the records. */
int rvp_profile_enable(int on);
/* restrict the timing to launches of one kernel (NULL or "" = all kernels) */
int rvp_profile_filter(const char *kernel);
int rvp_profile_collect(char *buf, size_t buflen, size_t *needed);
/* Process-wide tuning knobs (no counterpart in the reference; results never depend on them, only speed):
 *   "knn_cell_div"         k-NN grid edge = estimated k-th neighbour distance / this (default 2.4, [0.25, 16])
 *   "seg_cells_per_point"  dense cells a batched grid may give each point of a cloud before coarsening
 *                          (default 32, [1, 64]; at least 2^17 cells per cloud)
 *   "queries_per_wave"     consecutive queries one wave serves in the capped radius searches and the k-NN
 *                          SOR means (0 = from the query count: radius 64 for a few million queries and more,
 *                          fewer below; SOR means 2 up to a quarter million, at most 16; [0, 64])
 *   "work_steal"           how the batched k-NN SOR means hand out runs of queries: 0 strided over the launch's
 *                          waves; 1 persistent waves taking runs from per-XCD ranges by atomic counters and
 *                          stealing when theirs is drained; 2 per-XCD ranges strided without atomics
 *   "knn_start_cells"      half-width in cells of the first box of a k-NN search that has no bound from a
 *                          previous query (default 1: 3x3x3 cells; [1, 8]); the box then grows until certified
 *   "normals_per_wave"     consecutive queries one wave serves in the radius normals, which share staged
 *                          windows (0 = 8; [0, 64])
 * set() returns RVP_E_ARG for an unknown name or a value out of range. */
int rvp_tune_set(const char *name, double value);
int rvp_tune_get(const char *name, double *value);

/* ---- D1 / D2 (src/core/rvv_common.cpp:9-107) ---------------------------------------------------------
 * D1 get_dist_sq_rvv: out[i] = ((x-qx)^2 + (y-qy)^2) + (z-qz)^2, each op rounded once.
 * D2 get_inds_in_radius_rvv: (subset[i], d2) for subset points with d2 <= r2, in subset order; *n_out host. */
int rvp_dist_sq(const float *x, const float *y, const float *z, size_t n, float qx, float qy, float qz,
                float *out, rvp_stream_t s);
int rvp_inds_in_radius(const float *x, const float *y, const float *z, const int32_t *subset, size_t m, float qx,
                       float qy, float qz, float r2, int32_t *out_idx, float *out_d2, size_t *n_out,
                       rvp_stream_t s);

/* ---- H1/H2: SpatialHash::setInputCloud + build (src/search/spatial_hashing.cpp:76-163) ------------ */
int rvp_hash_create(rvp_hash **h);
int rvp_hash_destroy(rvp_hash *h);
/* Keeps a view of x/y/z (caller keeps them alive, as the reference does, spatial_hashing.cpp:77) and
 * builds the device grid: bbox with the 0.01*cell margin, cell keys, stable radix sort, cell table,
 * cell-ordered SoA copy. */
int rvp_hash_build(rvp_hash *h, const float *x, const float *y, const float *z, size_t n, float cell_size,
                   rvp_stream_t s);
/* host out: {minx,miny,minz,maxx,maxy,maxz} after the margin (spatial_hashing.cpp:141-146) */
int rvp_hash_bbox(const rvp_hash *h, float *bbox6);
/* host out: n points, occupied cells, grid dims (gx,gy,gz), internal cell size actually used */
int rvp_hash_info(const rvp_hash *h, size_t *n, size_t *occupied_cells, int *dims3, float *grid_cell);
/* device out: per point (original order) int32 cell coordinates ix,iy,iz (spatial_hashing.h:47-52) and the
 * reference's int64 key ix + iy*73856093 + iz*19349663 (spatial_hashing.h:43-45); either may be NULL. */
int rvp_hash_cell_coords(const rvp_hash *h, int32_t *ixyz, int64_t *keys, rvp_stream_t s);

/* ---- H3: SpatialHash::radiusSearch, batched (spatial_hashing.cpp:165-223) -------------------------
 * Query i's neighbours within `radius`, ascending (d2, index), truncated to max_nn.
 * max_nn > 0: rows of stride max_nn: idx_out[i*max_nn + j], d2_out likewise (either may be NULL),
 *             counts[i] = min(total_i, max_nn); total[i] (may be NULL) = untruncated count.
 * max_nn <= 0 (unlimited, the reference's default): use rvp_radius_count then rvp_radius_fill. */
int rvp_radius_search(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq,
                      float radius, int max_nn, int32_t *idx_out, float *d2_out, int32_t *counts,
                      int32_t *total, rvp_stream_t s);
/* counts[i] = number of neighbours of query i within radius (device out) */
int rvp_radius_count(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq,
                     float radius, int32_t *counts, rvp_stream_t s);
/* CSR fill: offsets[nq+1] (device, exclusive scan of counts, int64); writes each list sorted by (d2, idx). */
int rvp_radius_fill(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq,
                    float radius, const int64_t *offsets, int32_t *idx_out, float *d2_out, rvp_stream_t s);
/* exclusive scan helper: offsets[0..n] of counts[0..n) (device), host total */
int rvp_exclusive_scan(const int32_t *counts, size_t n, int64_t *offsets, int64_t *total_host, rvp_stream_t s);
/* 32-bit exclusive scan (device in/out, may alias; total_dev may be NULL): the single-pass scan the sorts and
 * compactions use; rvp_scan_wait_errors reports look-back waits that hit their spin limit (0 when healthy). */
int rvp_exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, rvp_stream_t s);
int rvp_scan_wait_errors(unsigned long long *count);
/* stable LSD radix sort of (key, value) pairs on the low `bits` bits of the key, 8-bit digits, one sweep per
 * digit: the sorted pairs end in (keys, vals), or in (keys_tmp, vals_tmp) when *in_tmp is set on return.
 * Device-sized form: n_dev (device, may be NULL) holds the count to sort (<= n), bits_dev (device, may be NULL)
 * the key width (<= bits); digits above it are not sorted on.  Neither is read on the host. */
int rvp_sort_pairs_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp, size_t n, int bits,
                       const uint32_t *n_dev, const int *bits_dev, int *in_tmp, rvp_stream_t s);
int rvp_sort_pairs_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_tmp, uint32_t *vals_tmp, size_t n, int bits,
                       const uint32_t *n_dev, const int *bits_dev, int *in_tmp, rvp_stream_t s);

/* ---- H4: radius_search_sc / radius_search_rvv, batched (src/search/radius_search.cpp:12-59) -------
 * Brute force: the first max_nn hits in index order; counts[i] = min(total, max_nn). */
int rvp_radius_bruteforce(const float *x, const float *y, const float *z, size_t n, const float *qx,
                          const float *qy, const float *qz, size_t nq, float radius, int max_nn,
                          int32_t *idx_out, float *d2_out, int32_t *counts, rvp_stream_t s);

/* ---- kNN over the hash: k nearest by (d2, idx), rows of stride k; rows shorter than k when n < k are
 * padded with idx -1, d2 +inf.  Query-minus-point order as sor_sc (statistical_outlier_removal.cpp:27-34). */
int rvp_knn(const rvp_hash *h, const float *qx, const float *qy, const float *qz, size_t nq, int k,
            int32_t *idx_out, float *d2_out, rvp_stream_t s);

/* ---- S1: sor_sc / sor_rvv (src/filters/statistical_outlier_removal.cpp:19-63, 68-190) --------------
 * Exact k-NN SOR.  out_aos (device, >= 3n floats) gets the kept points in input order; *n_out (host).
 * keep (device uint8[n]) and means (device float[n]) may be NULL. */
int rvp_sor_exact(const float *x, const float *y, const float *z, size_t n, int k, float alpha,
                  float *out_aos, size_t *n_out, uint8_t *keep, float *means, rvp_stream_t s);

/* ---- S2: sor_spatial_hash + filter_by_mean_dists (statistical_outlier_removal.cpp:246-270, 193-218) --
 * h must be built on the same x/y/z (it is the caller's SpatialHash). */
int rvp_sor_radius(const rvp_hash *h, int k, float alpha, float search_radius, float *out_aos,
                   size_t *n_out, uint8_t *keep, float *means, rvp_stream_t s);

/* The sequential fp32 sum the SOR statistics use (statistical_outlier_removal.cpp:43-52, 193-205), exact:
 * *out_dev = fl(...fl(fl(0 + x_0) + x_1)... + x_{n-1}) for non-negative terms x_i = v[i], or, when
 * squared_deviation != 0, x_i = fl(fl(v[i] - *mean_dev)^2).  Computed in parallel (binade-wise integer
 * composition, verified per chunk); out_dev and mean_dev are device pointers. */
int rvp_seq_sum(const float *v, size_t n, int squared_deviation, const float *mean_dev, float *out_dev,
                rvp_stream_t s);

/* ---- N1: normal_estimation_sc semantics (src/features/normal_estimation.cpp:72-112) ---------------
 * neighbourhood = k+1 nearest incl. self by (d2, idx); FP32 compensated covariance re-centred at the
 * query, closed-form FP32 eigen solve with FP64 fallback for near-degenerate spectra; flip toward vp
 * (normal_estimation.cpp:55-67).  nx/ny/nz device float[n]. */
int rvp_normals_knn(const float *x, const float *y, const float *z, size_t n, int k, float vpx, float vpy,
                    float vpz, float *nx, float *ny, float *nz, rvp_stream_t s);
/* The same with the neighbourhood centroid of every point (the mean of its k+1 neighbours, computed
 * re-centred at the point with compensated sums): cx/cy/cz device float[n], all NULL or none. */
int rvp_normals_knn_centroids(const float *x, const float *y, const float *z, size_t n, int k, float vpx, float vpy,
                              float vpz, float *nx, float *ny, float *nz, float *cx, float *cy, float *cz,
                              rvp_stream_t s);

/* ---- N2: normal_estimation_rvv(in, SpatialHash&, ...) (normal_estimation.cpp:189-214) --------------
 * neighbourhood = H3 list within radius (max_nn > 0 caps it, 0 = unlimited); fewer than 3 -> zero normal.
 * Queries are the hash's own points. */
int rvp_normals_radius(const rvp_hash *h, float radius, int max_nn, float vpx, float vpy, float vpz,
                       float *nx, float *ny, float *nz, rvp_stream_t s);
/* number of points that took the FP64 eigen fallback in the last normals call made by the calling host thread
 * (synchronises that call's stream; every stream has its own counter, so calls on other streams or threads do
 * not disturb it) */
int rvp_normals_fallback_count(size_t *count);

/* ---- V1: voxel_grid_downsamp_sc (src/filters/voxel_grid_downsamp.cpp:183-213) ---------------------
 * out_aos (device, >= 3n floats): centroids in lexicographic (vx,vy,vz) order; *n_out host. */
int rvp_voxel(const float *x, const float *y, const float *z, size_t n, float leaf, float *out_aos,
              size_t *n_out, rvp_stream_t s);

/* ---- R1: ransac_plane_sc / ransac_plane_rvv (src/segmentation/ransac_plane.cpp:55-111) -------------
 * Hypotheses from the glibc TYPE_3 rand() stream after srand(0) (3 draws per iteration), plane
 * coefficients as compute_plane_coefficients (:18-50), inlier counts on the GPU, adaptive stop replayed
 * on the host.  model (host float[4]); *inliers (host).  *iterations (host, may be NULL): iterations the
 * reference's loop runs (:68), which can exceed the final adaptive bound when it drops below the count. */
int rvp_ransac_plane(const float *x, const float *y, const float *z, size_t n, float dist_thresh,
                     int max_iters, float collinear_thresh, float probability, float *model, int *inliers,
                     int *iterations, rvp_stream_t s);
/* Counting core: counts[h] = #{i : fabsf(((a x + b y) + c z) + d) <= thr} for H models (device float[4H]). */
int rvp_plane_counts(const float *x, const float *y, const float *z, size_t n, const float *models, int H,
                     float dist_thresh, int32_t *counts, rvp_stream_t s);

/* ---- X1: extract_plane_inliers_outliers_rvv (ransac_plane.cpp:374-382, scalar order) -------------
 * Stable split into inliers / outliers (AoS, device); counts host.  Either output may be NULL.  One pass:
 * each tile's inlier rank comes from a decoupled look-back, its outlier rank is its index minus that. */
int rvp_plane_extract(const float *x, const float *y, const float *z, size_t n, const float *model_host,
                      float dist_thresh, float *inliers_aos, size_t *n_in, float *outliers_aos,
                      size_t *n_out, rvp_stream_t s);

/* Allocation-free variant of rvp_plane_extract (same arguments, same results).  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under graph capture, tight serving
 * loops).  It calls no hipMalloc / hipFree: its look-back status words, ticket counter and counts live in one
 * fixed array per device (0.5 MiB, held until the process ends), so n is at most 134,217,728 (larger n is
 * rejected with RVP_E_UNSUPPORTED).  Every call on a device shares that array: calls that overlap on one device
 * (other streams or threads) race and must be serialised by the caller. */
int rvp_plane_extract_noalloc(const float *x, const float *y, const float *z, size_t n, const float *model_host,
                              float dist_thresh, float *inliers_aos, size_t *n_in, float *outliers_aos,
                              size_t *n_out, rvp_stream_t s);

/* ---- R1 + X1 fused: the stage split of pipeline_export.cpp:457-498 --------------------------------
 * rvp_ransac_plane, then (only when a model was found, *inliers > 0) the one-pass split of the same points by
 * that model into inliers / outliers (AoS, device, index order; either may be NULL); counts host. */
int rvp_ransac_plane_extract(const float *x, const float *y, const float *z, size_t n, float dist_thresh,
                             int max_iters, float collinear_thresh, float probability, float *model, int *inliers,
                             int *iterations, float *inliers_aos, size_t *n_in, float *outliers_aos,
                             size_t *n_out, rvp_stream_t s);

/* ---- batches: many clouds, one launch per stage ---------------------------------------------------
 * The clouds lie back to back in x/y/z; offsets[nclouds+1] (device, uint32) delimit them; n_bound >= the
 * total number of points (buffers are sized by it; the true total is read from offsets on the device).
 * Every cloud's result equals what the single-cloud entry point returns for it.  These calls do not wait
 * on the host (rvp_batch_ransac_plane excepted): outputs (including out_offsets) are device memory, ordered
 * on the stream, so stages chain without a host round trip (the next stage takes out_offsets as its
 * offsets and the same n_bound). */

/* S1 per cloud (sor_sc, statistical_outlier_removal.cpp:19-63): kept points of cloud c at
 * out_aos[3*out_offsets[c] ..], in input order; keep / means (device, n_bound) may be NULL. */
int rvp_batch_sor_exact(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                        size_t n_bound, int k, float alpha, float *out_aos, uint32_t *out_offsets, uint8_t *keep,
                        float *means, rvp_stream_t s);
/* H3 per cloud (spatial_hashing.cpp:165-223, max_nn in [1, 64]): every point queries its own cloud with
 * radius[c] (device, per cloud) on a grid of edge cell[c] (NULL: radius); rows of stride max_nn indexed by
 * the point's batch position, neighbour indices local to the cloud. */
int rvp_batch_radius_search(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                            size_t n_bound, const float *radius, const float *cell, int max_nn, int32_t *idx_out,
                            float *d2_out, int32_t *counts, rvp_stream_t s);
/* N1 per cloud (normal_estimation.cpp:72-112 semantics, as rvp_normals_knn): outputs by batch position. */
int rvp_batch_normals_knn(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                          size_t n_bound, int k, float vpx, float vpy, float vpz, float *nx, float *ny, float *nz,
                          rvp_stream_t s);
/* rvp_batch_normals_knn (same outputs, bit for bit) reading each query's k + 1 nearest from the capped radius
 * lists of the same clouds (rvp_batch_radius_search's idx / d2 / counts, rows of stride list_cap, d2 may be
 * NULL) when its list holds at least k + 1 entries: the K nearest by (d2, idx) are then the list's first K.
 * Queries with shorter lists are searched.  No single reference function: it fuses the reference's
 * SpatialHash::radiusSearch (spatial_hashing.cpp:110-168) and normal_estimation_sc's k-NN
 * (normal_estimation.cpp:72-112) for pipelines that run both on one cloud. */
int rvp_batch_normals_knn_lists(const float *x, const float *y, const float *z, const uint32_t *offsets,
                                int nclouds, size_t n_bound, int k, const int32_t *list_idx, const float *list_d2,
                                const int32_t *list_counts, int list_cap, float vpx, float vpy, float vpz, float *nx,
                                float *ny, float *nz, rvp_stream_t s);
/* N2 per cloud with a cap (normal_estimation.cpp:189-214 neighbourhood, as rvp_normals_radius with
 * max_nn in [1, 64]): radius[c] / cell[c] device arrays (cell NULL: radius). */
int rvp_batch_normals_radius(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                             size_t n_bound, const float *radius, const float *cell, int max_nn, float vpx, float vpy,
                             float vpz, float *nx, float *ny, float *nz, rvp_stream_t s);
/* V1 per cloud (voxel_grid_downsamp.cpp:183-213, as rvp_voxel): leaf[c] device array; voxels of cloud c at
 * out_aos[3*out_offsets[c] ..] in lexicographic (vx, vy, vz) order.  status[c] (device) = 1 when cloud c's
 * voxel key range exceeds 2^48 (its output is then one centroid of the whole cloud), else 0. */
int rvp_batch_voxel(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                    size_t n_bound, const float *leaf, float *out_aos, uint32_t *out_offsets, uint32_t *status,
                    rvp_stream_t s);
/* R1 per cloud (ransac_plane.cpp:55-111, as rvp_ransac_plane): every hypothesis of every cloud is counted
 * in one launch, then the adaptive stop is replayed per cloud on the host.  Waits for the stream; outputs
 * are host arrays: models_out[4*nclouds], inliers_out[nclouds], iterations_out[nclouds] (may be NULL).
 * Unlike ransac_plane_sc it leaves the process rand() state alone. */
int rvp_batch_ransac_plane(const float *x, const float *y, const float *z, const uint32_t *offsets, int nclouds,
                           size_t n_bound, float dist_thresh, int max_iters, float collinear_thresh,
                           float probability, float *models_out, int *inliers_out, int *iterations_out,
                           rvp_stream_t s);

/* ---- PCD ingest (src/io/simple_pcd_loader.h:143-328) --------------------------------------------------
 * DATA ascii, binary and binary_compressed (LZF), x/y/z as 4-byte fields of any record layout, header rules
 * of the reference's parser.  info: point count from the header, data mode 0 ascii / 1 binary / 2
 * binary_compressed / -1 other.  read: host SoA; read_device: device SoA (binary records de-interleaved on
 * the GPU, compressed blocks copied straight in; returns after the copies).  *n receives the count even
 * when cap is too small (RVP_E_ARG).  write: as savePCD (:295-328), FIELDS x y z, ascii or binary. */
int rvp_pcd_info(const char *path, size_t *n_points, int *data_mode);
int rvp_pcd_read(const char *path, float *x, float *y, float *z, size_t cap, size_t *n);
int rvp_pcd_read_device(const char *path, float *x, float *y, float *z, size_t cap, size_t *n, rvp_stream_t s);
int rvp_pcd_write(const char *path, const float *x, const float *y, const float *z, size_t n, int binary);

/* ---- Euclidean clustering (src/segmentation/euclidean_clustering.cpp:113-189) -----------------------
 * Connected components of the graph linking points of the hash's cloud within `tolerance`
 * (((dx*dx + dy*dy) + dz*dz) <= fl(tol*tol), as radiusQueryUnvisited :98-109), which are the clusters the
 * reference's breadth-first search grows.  Clusters of min_size..max_size points, in the order of their
 * smallest index, each ascending: indices_out (device int32[n]) back to back, offsets_out (device
 * int64[n+1]) their bounds, *nclusters (host).  labels (device int32[n], may be NULL): every point's
 * component id, its smallest index.  Any grid edge works (the radius window adapts). */
int rvp_euclidean_cluster(const rvp_hash *h, float tolerance, int min_size, int max_size, int32_t *indices_out,
                          int64_t *offsets_out, size_t *nclusters, int32_t *labels, rvp_stream_t s);
/* union-find steps that hit their bound since load (0 when healthy) */
int rvp_cluster_errors(unsigned long long *count);

/* ---- streaming frames (no counterpart in the reference: the serving loop of the streaming workload) ---
 * A server runs `slots` batches at once, each on its own stream with buffers for slot_points points.  Frames
 * wait in one FIFO queue; a free slot takes as many as fit (continuous batching) and runs them through
 * voxel -> exact SOR -> radius normals (capped) -> RANSAC, each frame with its own results, equal to the
 * frame run alone.  hbm_budget bounds the device memory of the slots (512 B per slot point): fewer slots
 * run when it holds fewer.  submit copies the frame (AoS, host) and returns at once; wait blocks until the
 * frame's batch is done and hands over the result arrays (free them with rvp_result_free).  Times are
 * steady-clock seconds. */
typedef struct {
  float leaf;                           /* V1 leaf */
  int sor_k;                            /* S1 k, in [1, 63] */
  float sor_alpha;                      /* S1 std multiplier */
  float normal_radius;                  /* N2 radius */
  int normal_max_nn;                    /* N2 cap, in [1, 64] */
  float vp[3];                          /* F1 viewpoint */
  float ransac_thresh;                  /* R1 distance threshold */
  int ransac_iters;                     /* R1 max iterations */
  float ransac_prob;                    /* R1 probability */
} rvp_frame_params;

typedef struct {
  uint64_t ticket;
  size_t n_in, n_voxel, n_sor;          /* points in, voxels, points kept by SOR */
  float *points;                        /* kept points, AoS float[3 * n_sor] (pinned host memory owned by the
                                           result: release with rvp_result_free) */
  float *normals;                       /* their normals, float[3 * n_sor] */
  float model[4];                       /* RANSAC plane */
  int inliers, iterations;
  uint32_t voxel_status;                /* as rvp_batch_voxel's status */
  int batch_frames;                     /* frames in the batch this one ran in */
  double t_submit, t_start, t_done;
} rvp_frame_result;

typedef struct rvp_server rvp_server;
int rvp_server_create(rvp_server **srv, const rvp_frame_params *params, int slots, size_t slot_points,
                      size_t hbm_budget);
int rvp_server_slots(const rvp_server *srv, int *slots);
int rvp_server_submit(rvp_server *srv, const float *xyz_aos, size_t n, uint64_t *ticket);
int rvp_server_wait(rvp_server *srv, uint64_t ticket, rvp_frame_result *out);
int rvp_result_free(rvp_frame_result *r);
int rvp_server_destroy(rvp_server *srv);

#ifdef __cplusplus
}
#endif
#endif
