/*
 * oracle.c -- plain-C restatement of RVPoint's scalar neighbour-search path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compile with -ffp-contract=off.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(__FP_FAST_FMA) && !defined(RVP_ORACLE_ALLOW_FMA)
/* contraction must be off: each fp32 op of the reference rounds once */
#endif

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */

/* D1: get_dist_sq_rvv scalar branch (rvv_common.cpp:36-43): ((dx*dx + dy*dy) + dz*dz), dx = p - q. */
static inline float d1(float px, float py, float pz, float qx, float qy, float qz) {
    float dx = px - qx;
    float dy = py - qy;
    float dz = pz - qz;
    float a = dx * dx;
    float b = dy * dy;
    float c = dz * dz;
    float s = a + b;
    return s + c;
}

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* (d2, idx) packed so that unsigned order == std::pair<float,int> order for d2 >= +0 */
static inline uint64_t pack(float d2, int32_t idx) { return ((uint64_t)fbits(d2) << 32) | (uint32_t)idx; }

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* partial selection: the k smallest of v[0..n) moved to v[0..k), sorted ascending */
static void select_sorted_u64(uint64_t *v, size_t n, size_t k) {
    if (k >= n) { qsort(v, n, sizeof(uint64_t), cmp_u64); return; }
    size_t lo = 0, hi = n - 1;
    while (hi > lo) {
        /* median of three pivot */
        size_t mid = lo + (hi - lo) / 2;
        uint64_t a = v[lo], b = v[mid], c = v[hi];
        uint64_t piv = (a < b) ? ((b < c) ? b : ((a < c) ? c : a)) : ((a < c) ? a : ((b < c) ? c : b));
        size_t i = lo, j = hi;
        while (i <= j) {
            while (v[i] < piv) i++;
            while (v[j] > piv) j--;
            if (i <= j) {
                uint64_t t = v[i]; v[i] = v[j]; v[j] = t;
                i++;
                if (j == 0) break;
                j--;
            }
        }
        if (k - 1 <= j) hi = j;
        else if (k - 1 >= i) lo = i;
        else break;
    }
    qsort(v, k, sizeof(uint64_t), cmp_u64);
}

/* ------------------------------------------------------------------ */
/* H1 / H2                                                             */
/* ------------------------------------------------------------------ */

/* spatial_hashing.cpp:76-146 (scalar branch :131-138, margin :141-146). */
/* get_dist_sq_rvv, scalar branch (rvv_common.cpp:36-43): out[i] = D1(p_i, q). */
void orc_dist_sq(const float *x, const float *y, const float *z, size_t n, float qx, float qy, float qz, float *out) {
    for (size_t i = 0; i < n; ++i) out[i] = d1(x[i], y[i], z[i], qx, qy, qz);
}

/* get_inds_in_radius_rvv, scalar branch (rvv_common.cpp:55-68): gather the subset, D1, and append
 * (subset[i], d2) for d2 <= r2 in subset order after the `have` entries already in the outputs; returns the new
 * length. */
size_t orc_inds_in_radius(const float *x, const float *y, const float *z, const int32_t *subset, size_t m,
                          float qx, float qy, float qz, float r2, int32_t *out_idx, float *out_d2, size_t have) {
    for (size_t i = 0; i < m; ++i) {
        const int32_t j = subset[i];
        const float d = d1(x[j], y[j], z[j], qx, qy, qz);
        if (d <= r2) { out_idx[have] = j; out_d2[have] = d; ++have; }
    }
    return have;
}

void orc_bbox(const float *x, const float *y, const float *z, size_t n, float cell, float out[6]) {
    float mnx = FLT_MAX, mny = FLT_MAX, mnz = FLT_MAX;
    float mxx = -FLT_MAX, mxy = -FLT_MAX, mxz = -FLT_MAX;
    if (n == 0) { memset(out, 0, 6 * sizeof(float)); return; }
    for (size_t j = 0; j < n; ++j) {
        if (x[j] < mnx) mnx = x[j];
        if (x[j] > mxx) mxx = x[j];
        if (y[j] < mny) mny = y[j];
        if (y[j] > mxy) mxy = y[j];
        if (z[j] < mnz) mnz = z[j];
        if (z[j] > mxz) mxz = z[j];
    }
    float eps_scale = 0.01f;
    float m = cell * eps_scale;
    mnx -= m; mny -= m; mnz -= m;
    mxx += m; mxy += m; mxz += m;
    out[0] = mnx; out[1] = mny; out[2] = mnz;
    out[3] = mxx; out[4] = mxy; out[5] = mxz;
}

static inline int cell_index(float v, float mn, float cell) {
    float t = (v - mn) / cell;              /* IEEE division, spatial_hashing.h:49 */
    return (int)floorf(t);
}

static inline int64_t hash_key(int ix, int iy, int iz) {   /* spatial_hashing.h:43-45 */
    return (int64_t)ix + (int64_t)iy * 73856093LL + (int64_t)iz * 19349663LL;
}

void orc_cell_coords(const float *x, const float *y, const float *z, size_t n, float cell,
                     int32_t *ixyz, int64_t *keys) {
    float bb[6];
    orc_bbox(x, y, z, n, cell, bb);
    for (size_t i = 0; i < n; ++i) {
        int ix = cell_index(x[i], bb[0], cell);
        int iy = cell_index(y[i], bb[1], cell);
        int iz = cell_index(z[i], bb[2], cell);
        ixyz[3 * i + 0] = ix; ixyz[3 * i + 1] = iy; ixyz[3 * i + 2] = iz;
        if (keys) keys[i] = hash_key(ix, iy, iz);
    }
}

/* ------------------------------------------------------------------ */
/* H3: hash grid as a key-sorted (key, idx) table                      */
/* ------------------------------------------------------------------ */

typedef struct { int64_t key; int32_t idx; } kv_t;

static int cmp_kv(const void *a, const void *b) {
    const kv_t *p = (const kv_t *)a, *q = (const kv_t *)b;
    if (p->key != q->key) return (p->key > q->key) - (p->key < q->key);
    return (p->idx > q->idx) - (p->idx < q->idx);     /* push_back order: ascending i (spatial_hashing.cpp:157-162) */
}

typedef struct { kv_t *t; size_t n; float bb[6]; float cell; } orc_grid;

static void grid_build(orc_grid *g, const float *x, const float *y, const float *z, size_t n, float cell) {
    g->n = n; g->cell = cell;
    g->t = (kv_t *)malloc((n ? n : 1) * sizeof(kv_t));
    orc_bbox(x, y, z, n, cell, g->bb);
    for (size_t i = 0; i < n; ++i) {
        int ix = cell_index(x[i], g->bb[0], cell);
        int iy = cell_index(y[i], g->bb[1], cell);
        int iz = cell_index(z[i], g->bb[2], cell);
        g->t[i].key = hash_key(ix, iy, iz);
        g->t[i].idx = (int32_t)i;
    }
    qsort(g->t, n, sizeof(kv_t), cmp_kv);
}

static size_t lower_key(const orc_grid *g, int64_t key) {
    size_t lo = 0, hi = g->n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (g->t[mid].key < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* radiusSearch (spatial_hashing.cpp:110-168): window of (2w+1)^3 buckets, w = ceil(r/c)+1; per bucket the
 * gather-filter of get_inds_in_radius (rvv_common.cpp:55-68, p - q order); sort by (d2, idx); truncate. */
static size_t grid_query(const orc_grid *g, const float *x, const float *y, const float *z,
                         float qx, float qy, float qz, float radius, uint64_t **buf, size_t *cap) {
    size_t m = 0;
    if (g->n == 0) return 0;
    float r2 = radius * radius;
    int qix = cell_index(qx, g->bb[0], g->cell);
    int qiy = cell_index(qy, g->bb[1], g->cell);
    int qiz = cell_index(qz, g->bb[2], g->cell);
    int w = (int)ceilf(radius / g->cell) + 1;
    for (int dx = -w; dx <= w; ++dx)
        for (int dy = -w; dy <= w; ++dy)
            for (int dz = -w; dz <= w; ++dz) {
                int64_t key = hash_key(qix + dx, qiy + dy, qiz + dz);
                size_t s = lower_key(g, key);
                for (size_t e = s; e < g->n && g->t[e].key == key; ++e) {
                    int32_t j = g->t[e].idx;
                    float d2 = d1(x[j], y[j], z[j], qx, qy, qz);
                    if (d2 <= r2) {
                        if (m == *cap) { *cap = *cap ? 2 * *cap : 256; *buf = (uint64_t *)realloc(*buf, *cap * 8); }
                        (*buf)[m++] = pack(d2, j);
                    }
                }
            }
    qsort(*buf, m, sizeof(uint64_t), cmp_u64);
    return m;
}

static void write_rows(const uint64_t *v, size_t m, int max_nn, size_t row, size_t stride,
                       int32_t *idx_out, float *d2_out, int32_t *counts, int32_t *total) {
    size_t keep = (max_nn > 0 && m > (size_t)max_nn) ? (size_t)max_nn : m;
    if (total) total[row] = (int32_t)m;
    if (counts) counts[row] = (int32_t)keep;
    if (!idx_out && !d2_out) return;
    for (size_t j = 0; j < keep && j < stride; ++j) {
        if (idx_out) idx_out[row * stride + j] = (int32_t)(v[j] & 0xffffffffu);
        if (d2_out) d2_out[row * stride + j] = bitsf((uint32_t)(v[j] >> 32));
    }
}

void orc_hash_radius(const float *x, const float *y, const float *z, size_t n, float cell,
                     const float *qx, const float *qy, const float *qz, size_t nq,
                     float radius, int max_nn, size_t row_stride,
                     int32_t *idx_out, float *d2_out, int32_t *counts, int32_t *total) {
    orc_grid g;
    grid_build(&g, x, y, z, n, cell);
    uint64_t *buf = NULL; size_t cap = 0;
    for (size_t q = 0; q < nq; ++q) {
        size_t m = grid_query(&g, x, y, z, qx[q], qy[q], qz[q], radius, &buf, &cap);
        write_rows(buf, m, max_nn, q, row_stride, idx_out, d2_out, counts, total);
    }
    free(buf);
    free(g.t);
}

void orc_radius_all(const float *x, const float *y, const float *z, size_t n,
                    const float *qx, const float *qy, const float *qz, size_t nq,
                    float radius, int max_nn, size_t row_stride,
                    int32_t *idx_out, float *d2_out, int32_t *counts, int32_t *total) {
    float r2 = radius * radius;
    uint64_t *buf = (uint64_t *)malloc((n ? n : 1) * 8);
    for (size_t q = 0; q < nq; ++q) {
        size_t m = 0;
        for (size_t j = 0; j < n; ++j) {
            float d2 = d1(x[j], y[j], z[j], qx[q], qy[q], qz[q]);
            if (d2 <= r2) buf[m++] = pack(d2, (int32_t)j);
        }
        qsort(buf, m, 8, cmp_u64);
        write_rows(buf, m, max_nn, q, row_stride, idx_out, d2_out, counts, total);
    }
    free(buf);
}

/* H4: radius_search_sc (radius_search.cpp:12-33). */
size_t orc_radius_bruteforce(const float *x, const float *y, const float *z, size_t n,
                             float qx, float qy, float qz, float radius,
                             int32_t *indices, float *dists, int max_nn) {
    size_t count = 0;
    float r2 = radius * radius;
    for (size_t i = 0; i < n; ++i) {
        float d2 = d1(x[i], y[i], z[i], qx, qy, qz);
        if (d2 <= r2) {
            if (count < (size_t)max_nn) {
                indices[count] = (int32_t)i;
                if (dists) dists[count] = d2;
            }
            count++;
        }
    }
    return (count > (size_t)max_nn) ? (size_t)max_nn : count;
}

/* ------------------------------------------------------------------ */
/* kNN                                                                 */
/* ------------------------------------------------------------------ */

void orc_knn(const float *x, const float *y, const float *z, size_t n,
             const float *qx, const float *qy, const float *qz, size_t nq,
             int k, int32_t *idx_out, float *d2_out) {
    uint64_t *buf = (uint64_t *)malloc((n ? n : 1) * 8);
    size_t kk = (size_t)k;
    for (size_t q = 0; q < nq; ++q) {
        for (size_t j = 0; j < n; ++j)
            buf[j] = pack(d1(qx[q], qy[q], qz[q], x[j], y[j], z[j]), (int32_t)j);  /* q - p, as sor_sc */
        size_t take = kk < n ? kk : n;
        select_sorted_u64(buf, n, take);
        for (size_t j = 0; j < kk; ++j) {
            if (j < take) {
                idx_out[q * kk + j] = (int32_t)(buf[j] & 0xffffffffu);
                d2_out[q * kk + j] = bitsf((uint32_t)(buf[j] >> 32));
            } else {
                idx_out[q * kk + j] = -1;
                d2_out[q * kk + j] = INFINITY;
            }
        }
    }
    free(buf);
}

/* ------------------------------------------------------------------ */
/* SOR                                                                 */
/* ------------------------------------------------------------------ */

/* filter_by_mean_dists (statistical_outlier_removal.cpp:193-218) == sor_sc steps 2-3 (:43-62). */
size_t orc_sor_filter(const float *x, const float *y, const float *z, size_t n,
                      const float *means, float alpha, float *out, uint8_t *keep, float stats[3]) {
    if (n == 0) return 0;
    float global_sum = 0.0f;
    for (size_t i = 0; i < n; ++i) global_sum += means[i];
    float nf = (float)n;
    float global_mean = global_sum / nf;
    float variance_sum = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float diff = means[i] - global_mean;
        float sq = diff * diff;
        variance_sum += sq;
    }
    float global_std = sqrtf(variance_sum / nf);
    float as = alpha * global_std;
    float thresh = global_mean + as;
    if (stats) { stats[0] = global_mean; stats[1] = global_std; stats[2] = thresh; }
    size_t count = 0;
    for (size_t i = 0; i < n; ++i) {
        int k = means[i] <= thresh;
        if (keep) keep[i] = (uint8_t)k;
        if (k) {
            if (out) { out[3 * count] = x[i]; out[3 * count + 1] = y[i]; out[3 * count + 2] = z[i]; }
            count++;
        }
    }
    return count;
}

float orc_seq_sum(const float *v, size_t n, int squared_dev, float mean) {
    float s = 0.0f;
    for (size_t i = 0; i < n; ++i) {
        float x = v[i];
        if (squared_dev) {
            float d = x - mean;
            x = d * d;
        }
        s += x;
    }
    return s;
}

/* sor_sc (statistical_outlier_removal.cpp:19-63): exact k+1 smallest of all d2 (incl. self), mean of sqrt of
 * entries 1..k summed in ascending order, divided by k. */
size_t orc_sor_sc(const float *aos, size_t n, int k, float alpha, float *out, float *means) {
    if (n == 0) return 0;
    float *md = (float *)malloc(n * sizeof(float));
    uint64_t *buf = (uint64_t *)malloc(n * 8);
    float *x = (float *)malloc(n * 4), *y = (float *)malloc(n * 4), *z = (float *)malloc(n * 4);
    for (size_t i = 0; i < n; ++i) { x[i] = aos[3 * i]; y[i] = aos[3 * i + 1]; z[i] = aos[3 * i + 2]; }
    size_t take = (size_t)k + 1 < n ? (size_t)k + 1 : n;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j) buf[j] = pack(d1(x[i], y[i], z[i], x[j], y[j], z[j]), (int32_t)j);
        select_sorted_u64(buf, n, take);
        float sum = 0.0f;
        for (int j = 1; j <= k && (size_t)j < take; ++j) sum += sqrtf(bitsf((uint32_t)(buf[j] >> 32)));
        md[i] = sum / (float)k;
    }
    size_t c = orc_sor_filter(x, y, z, n, md, alpha, out, NULL, NULL);
    if (means) memcpy(means, md, n * sizeof(float));
    free(md); free(buf); free(x); free(y); free(z);
    return c;
}

/* sor_sc with the per-point means computed by `threads` OpenMP threads (each point's k nearest and its mean
 * are the same arithmetic as orc_sor_sc; the global mean and deviation sums stay sequential in index order,
 * statistical_outlier_removal.cpp:43-52), so the outputs are bit-identical to orc_sor_sc.  The loop over
 * independent points with thread-private buffers follows the reference's own OpenMP pattern
 * (tests/unit/test_openmp_rvv_scaling.cpp:61-104). */
size_t orc_sor_sc_mt(const float *aos, size_t n, int k, float alpha, float *out, float *means, int threads) {
    if (n == 0) return 0;
    float *md = (float *)malloc(n * sizeof(float));
    float *x = (float *)malloc(n * 4), *y = (float *)malloc(n * 4), *z = (float *)malloc(n * 4);
    for (size_t i = 0; i < n; ++i) { x[i] = aos[3 * i]; y[i] = aos[3 * i + 1]; z[i] = aos[3 * i + 2]; }
    const size_t take = (size_t)k + 1 < n ? (size_t)k + 1 : n;
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        uint64_t *buf = (uint64_t *)malloc(n * 8);
#pragma omp for schedule(dynamic, 64)
        for (long long ii = 0; ii < (long long)n; ++ii) {
            const size_t i = (size_t)ii;
            for (size_t j = 0; j < n; ++j) buf[j] = pack(d1(x[i], y[i], z[i], x[j], y[j], z[j]), (int32_t)j);
            select_sorted_u64(buf, n, take);
            float sum = 0.0f;
            for (int j = 1; j <= k && (size_t)j < take; ++j) sum += sqrtf(bitsf((uint32_t)(buf[j] >> 32)));
            md[i] = sum / (float)k;
        }
        free(buf);
    }
    size_t c = orc_sor_filter(x, y, z, n, md, alpha, out, NULL, NULL);
    if (means) memcpy(means, md, n * sizeof(float));
    free(md); free(x); free(y); free(z);
    return c;
}

/* sor_spatial_hash (statistical_outlier_removal.cpp:246-270). */
size_t orc_sor_radius(const float *x, const float *y, const float *z, size_t n, float cell,
                      int k, float alpha, float search_radius, float *out, float *means) {
    if (n == 0) return 0;
    orc_grid g;
    grid_build(&g, x, y, z, n, cell);
    float *md = (float *)malloc(n * sizeof(float));
    uint64_t *buf = NULL; size_t cap = 0;
    for (size_t i = 0; i < n; ++i) {
        size_t m = grid_query(&g, x, y, z, x[i], y[i], z[i], search_radius, &buf, &cap);
        if (m > (size_t)k + 1) m = (size_t)k + 1;                  /* max_nn = k+1 (:255) */
        if (m > 1) {
            float sum = 0.0f;
            int valid_k = (k < (int)m - 1) ? k : (int)m - 1;
            for (int j = 1; j <= valid_k; ++j) sum += sqrtf(bitsf((uint32_t)(buf[j] >> 32)));
            md[i] = sum / (float)valid_k;
        } else {
            md[i] = search_radius;
        }
    }
    size_t c = orc_sor_filter(x, y, z, n, md, alpha, out, NULL, NULL);
    if (means) memcpy(means, md, n * sizeof(float));
    free(md); free(buf); free(g.t);
    return c;
}

/* ------------------------------------------------------------------ */
/* V1                                                                  */
/* ------------------------------------------------------------------ */

typedef struct { int32_t v[3]; int32_t i; } vkey_t;

static int cmp_vkey(const void *a, const void *b) {
    const vkey_t *p = (const vkey_t *)a, *q = (const vkey_t *)b;
    for (int d = 0; d < 3; ++d)
        if (p->v[d] != q->v[d]) return (p->v[d] > q->v[d]) - (p->v[d] < q->v[d]);
    return (p->i > q->i) - (p->i < q->i);
}

/* voxel_grid_downsamp_sc (voxel_grid_downsamp.cpp:183-213): std::map order == lexicographic (vx,vy,vz);
 * per-voxel sums accumulate in input order; centroid = sum * (1.0f / count). */
size_t orc_voxel_sc(const float *aos, size_t n, float leaf, float *out) {
    if (n == 0) return 0;
    vkey_t *t = (vkey_t *)malloc(n * sizeof(vkey_t));
    float inv_leaf = 1.0f / leaf;
    for (size_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            float s = aos[3 * i + d] * inv_leaf;
            t[i].v[d] = (int)floorf(s);
        }
        t[i].i = (int32_t)i;
    }
    qsort(t, n, sizeof(vkey_t), cmp_vkey);
    size_t count = 0, s = 0;
    while (s < n) {
        size_t e = s;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        int c = 0;
        while (e < n && t[e].v[0] == t[s].v[0] && t[e].v[1] == t[s].v[1] && t[e].v[2] == t[s].v[2]) {
            size_t i = (size_t)t[e].i;
            sx += aos[3 * i]; sy += aos[3 * i + 1]; sz += aos[3 * i + 2];
            c++; e++;
        }
        float f = 1.0f / (float)c;
        out[3 * count] = sx * f; out[3 * count + 1] = sy * f; out[3 * count + 2] = sz * f;
        count++;
        s = e;
    }
    free(t);
    return count;
}

/* ------------------------------------------------------------------ */
/* Normals (FP64 oracle) + E1/F1 restatements                          */
/* ------------------------------------------------------------------ */

/* cyclic Jacobi on a symmetric 3x3 (double), to convergence */
static void jacobi3(double A[3][3], double V[3][3]) {
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) V[i][j] = (i == j);
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
        double dia = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
        if (off == 0.0 || off <= 1e-300 || off < 1e-18 * dia) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {         /* A <- A J */
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {         /* A <- J^T A */
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
}

void orc_normal_from_points_f64(const double *px, const double *py, const double *pz, size_t m,
                                double n_out[3], double evals[3]) {
    double cx = 0, cy = 0, cz = 0;
    for (size_t j = 0; j < m; ++j) { cx += px[j]; cy += py[j]; cz += pz[j]; }
    cx /= (double)m; cy /= (double)m; cz /= (double)m;
    double A[3][3] = {{0}};
    for (size_t j = 0; j < m; ++j) {
        double dx = px[j] - cx, dy = py[j] - cy, dz = pz[j] - cz;
        A[0][0] += dx * dx; A[0][1] += dx * dy; A[0][2] += dx * dz;
        A[1][1] += dy * dy; A[1][2] += dy * dz; A[2][2] += dz * dz;
    }
    A[1][0] = A[0][1]; A[2][0] = A[0][2]; A[2][1] = A[1][2];
    double V[3][3];
    jacobi3(A, V);
    int order[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a)
        for (int b = a + 1; b < 3; ++b)
            if (A[order[b]][order[b]] < A[order[a]][order[a]]) { int t = order[a]; order[a] = order[b]; order[b] = t; }
    int mi = order[0];
    double nx = V[0][mi], ny = V[1][mi], nz = V[2][mi];
    double nn = sqrt(nx * nx + ny * ny + nz * nz);
    n_out[0] = nx / nn; n_out[1] = ny / nn; n_out[2] = nz / nn;
    if (evals) { evals[0] = A[order[0]][order[0]]; evals[1] = A[order[1]][order[1]]; evals[2] = A[order[2]][order[2]]; }
}

static void flip_f64(double px, double py, double pz, const float vp[3], double n[3]) {
    double vx = (double)vp[0] - px, vy = (double)vp[1] - py, vz = (double)vp[2] - pz;
    double dot = vx * n[0] + vy * n[1] + vz * n[2];
    if (dot < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
}

void orc_normals_knn_f64(const float *x, const float *y, const float *z, size_t n, int k,
                         const float vp[3], double *nx, double *ny, double *nz) {
    if (n == 0) return;
    const size_t K = (size_t)k + 1 < n ? (size_t)k + 1 : n;
    /* points are independent: OpenMP over them with thread-private buffers (results do not depend on it) */
#pragma omp parallel
    {
        uint64_t *buf = (uint64_t *)malloc(n * 8);
        double *px = (double *)malloc(K * 8), *py = (double *)malloc(K * 8), *pz = (double *)malloc(K * 8);
#pragma omp for schedule(dynamic, 64)
        for (long long ii = 0; ii < (long long)n; ++ii) {
            const size_t i = (size_t)ii;
            for (size_t j = 0; j < n; ++j) buf[j] = pack(d1(x[i], y[i], z[i], x[j], y[j], z[j]), (int32_t)j);
            select_sorted_u64(buf, n, K);
            for (size_t j = 0; j < K; ++j) {
                int32_t id = (int32_t)(buf[j] & 0xffffffffu);
                px[j] = x[id]; py[j] = y[id]; pz[j] = z[id];
            }
            double nv[3];
            orc_normal_from_points_f64(px, py, pz, K, nv, NULL);
            flip_f64(x[i], y[i], z[i], vp, nv);
            nx[i] = nv[0]; ny[i] = nv[1]; nz[i] = nv[2];
        }
        free(buf); free(px); free(py); free(pz);
    }
}

void orc_normals_radius_f64(const float *x, const float *y, const float *z, size_t n, float cell,
                            float radius, int max_nn, const float vp[3],
                            double *nx, double *ny, double *nz) {
    if (n == 0) return;
    orc_grid g;
    grid_build(&g, x, y, z, n, cell);
#pragma omp parallel
    {
        uint64_t *buf = NULL; size_t cap = 0;
        size_t pcap = 64;
        double *px = (double *)malloc(pcap * 8), *py = (double *)malloc(pcap * 8), *pz = (double *)malloc(pcap * 8);
#pragma omp for schedule(dynamic, 256)
        for (long long ii = 0; ii < (long long)n; ++ii) {
            const size_t i = (size_t)ii;
            size_t m = grid_query(&g, x, y, z, x[i], y[i], z[i], radius, &buf, &cap);
            if (max_nn > 0 && m > (size_t)max_nn) m = (size_t)max_nn;
            if (m < 3) { nx[i] = ny[i] = nz[i] = 0.0; continue; }
            if (m > pcap) {
                pcap = m;
                px = (double *)realloc(px, pcap * 8); py = (double *)realloc(py, pcap * 8); pz = (double *)realloc(pz, pcap * 8);
            }
            for (size_t j = 0; j < m; ++j) {
                int32_t id = (int32_t)(buf[j] & 0xffffffffu);
                px[j] = x[id]; py[j] = y[id]; pz[j] = z[id];
            }
            double nv[3];
            orc_normal_from_points_f64(px, py, pz, m, nv, NULL);
            flip_f64(x[i], y[i], z[i], vp, nv);
            nx[i] = nv[0]; ny[i] = nv[1]; nz[i] = nv[2];
        }
        free(buf); free(px); free(py); free(pz);
    }
    free(g.t);
}

/* simple_eigen3x3_smallest (normal_estimation.cpp:16-52), op for op. */
void orc_eigen_e1(const float cov[9], int iters, float n_out[3]) {
    float A[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) A[i][j] = cov[3 * i + j];
    float V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int iter = 0; iter < iters; ++iter) {
        int p = 0, q = 1;
        float max_off = fabsf(A[0][1]);
        if (fabsf(A[0][2]) > max_off) { p = 0; q = 2; max_off = fabsf(A[0][2]); }
        if (fabsf(A[1][2]) > max_off) { p = 1; q = 2; }
        float phi = 0.5f * atan2f(2 * A[p][q], A[p][p] - A[q][q]);
        float c = cosf(phi);
        float s = sinf(phi);
        float app = A[p][p], aqq = A[q][q], apq = A[p][q];
        A[p][p] = c * c * app - 2 * s * c * apq + s * s * aqq;
        A[q][q] = s * s * app + 2 * s * c * apq + c * c * aqq;
        A[p][q] = 0;
        for (int k = 0; k < 3; ++k) {
            float vip = V[k][p];
            float viq = V[k][q];
            V[k][p] = c * vip - s * viq;
            V[k][q] = s * vip + c * viq;
        }
    }
    int min_idx = 0;
    if (A[1][1] < A[min_idx][min_idx]) min_idx = 1;
    if (A[2][2] < A[min_idx][min_idx]) min_idx = 2;
    n_out[0] = V[0][min_idx]; n_out[1] = V[1][min_idx]; n_out[2] = V[2][min_idx];
}

/* flipNormalTowardsViewpoint (normal_estimation.cpp:55-67). */
void orc_flip(float px, float py, float pz, float vx_, float vy_, float vz_, float n[3]) {
    float vx = vx_ - px, vy = vy_ - py, vz = vz_ - pz;
    float dot = vx * n[0] + vy * n[1] + vz * n[2];
    if (dot < 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
}

/* normal_estimation_rvv(in, SpatialHash, ...) (normal_estimation.cpp:189-214) over `threads` OpenMP threads:
 * per point the hash radius list (no cap), compute_covariance_rvv's sequential centroid and raw covariance
 * sums in list order (:117-142), the four-step eigen step E1 (:16-52) and the viewpoint flip (:55-67); fewer
 * than 3 neighbours give a zero normal.  Points are independent, so the result does not depend on threads. */
void orc_normals_hash_e1_mt(const float *x, const float *y, const float *z, size_t n, float cell, float radius,
                            const float vp[3], int eigen_iters, float *nx, float *ny, float *nz, int threads) {
    if (n == 0) return;
    orc_grid g;
    grid_build(&g, x, y, z, n, cell);
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
    {
        uint64_t *buf = NULL;
        size_t cap = 0;
#pragma omp for schedule(dynamic, 256)
        for (long long ii = 0; ii < (long long)n; ++ii) {
            const size_t i = (size_t)ii;
            const size_t m = grid_query(&g, x, y, z, x[i], y[i], z[i], radius, &buf, &cap);
            if (m < 3) { nx[i] = ny[i] = nz[i] = 0.0f; continue; }
            float cx = 0, cy = 0, cz = 0;
            for (size_t j = 0; j < m; ++j) {
                const uint32_t q = (uint32_t)(buf[j] & 0xffffffffu);
                cx += x[q]; cy += y[q]; cz += z[q];
            }
            const float inv_n = 1.0f / (float)m;
            cx *= inv_n; cy *= inv_n; cz *= inv_n;
            float c00 = 0, c01 = 0, c02 = 0, c11 = 0, c12 = 0, c22 = 0;
            for (size_t j = 0; j < m; ++j) {
                const uint32_t q = (uint32_t)(buf[j] & 0xffffffffu);
                const float dx = x[q] - cx, dy = y[q] - cy, dz = z[q] - cz;
                c00 += dx * dx; c01 += dx * dy; c02 += dx * dz;
                c11 += dy * dy; c12 += dy * dz;
                c22 += dz * dz;
            }
            const float cov[9] = {c00, c01, c02, c01, c11, c12, c02, c12, c22};
            float nv[3];
            orc_eigen_e1(cov, eigen_iters, nv);
            orc_flip(x[i], y[i], z[i], vp[0], vp[1], vp[2], nv);
            nx[i] = nv[0]; ny[i] = nv[1]; nz[i] = nv[2];
        }
        free(buf);
    }
    free(g.t);
}

/* ------------------------------------------------------------------ */
/* glibc TYPE_3 rand() (published algorithm: additive feedback r[i] = r[i-3] + r[i-31], output >> 1,
 * state seeded by the Park-Miller LCG 16807 mod 2^31-1, 310 outputs discarded).                      */
/* ------------------------------------------------------------------ */

void orc_srand(orc_rand_state *s, unsigned seed) {
    int32_t r[34];
    if (seed == 0) seed = 1;
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
        int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int64_t w = 16807 * lo - 2836 * hi;
        if (w < 0) w += 2147483647;
        r[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    memcpy(s->r, r, sizeof(r));
    s->pos = 0;                       /* ring index of r[i-34] for the next i */
    for (int i = 34; i < 344; ++i) (void)orc_rand(s);
}

int32_t orc_rand(orc_rand_state *s) {
    /* ring of the last 34 words: r[i] = r[i-31] + r[i-3] */
    int p = s->pos;                             /* slot holding r[i-34] */
    uint32_t a = (uint32_t)s->r[(p + 3) % 34];  /* r[i-31] */
    uint32_t b = (uint32_t)s->r[(p + 31) % 34]; /* r[i-3]  */
    uint32_t v = a + b;
    s->r[p] = (int32_t)v;
    s->pos = (p + 1) % 34;
    return (int32_t)(v >> 1);
}

/* ------------------------------------------------------------------ */
/* RANSAC                                                              */
/* ------------------------------------------------------------------ */

/* compute_plane_coefficients (ransac_plane.cpp:18-50). */
int orc_plane_model(const float p1[3], const float p2[3], const float p3[3], float collinear, float model[4]) {
    float v1x = p2[0] - p1[0], v1y = p2[1] - p1[1], v1z = p2[2] - p1[2];
    float v2x = p3[0] - p1[0], v2y = p3[1] - p1[1], v2z = p3[2] - p1[2];
    float a = v1y * v2z - v1z * v2y;
    float b = v1z * v2x - v1x * v2z;
    float c = v1x * v2y - v1y * v2x;
    float norm = sqrtf(a * a + b * b + c * c);
    if (norm < collinear) return 0;
    a /= norm; b /= norm; c /= norm;
    float d = -(a * p1[0] + b * p1[1] + c * p1[2]);
    model[0] = a; model[1] = b; model[2] = c; model[3] = d;
    return 1;
}

int orc_plane_count(const float *x, const float *y, const float *z, size_t n, const float model[4], float thr) {
    int cnt = 0;
    for (size_t i = 0; i < n; ++i) {
        float dist = fabsf(model[0] * x[i] + model[1] * y[i] + model[2] * z[i] + model[3]);
        if (dist <= thr) cnt++;
    }
    return cnt;
}

/* ransac_plane_sc (ransac_plane.cpp:55-111). */
int orc_ransac_plane(const float *x, const float *y, const float *z, size_t n, float thr, int max_iters,
                     float collinear, float probability, float model[4], int *iters_run) {
    if (iters_run) *iters_run = 0;
    if (n < 3) return 0;
    orc_rand_state st;
    orc_srand(&st, 0);
    int best = 0;
    float best_model[4] = {0, 0, 0, 0};
    int k_iters = max_iters;
    double pc = (double)probability;
    if (pc < 0.5) pc = 0.5;
    if (pc > 0.9999) pc = 0.9999;
    const double log_p = log(1.0 - pc);
    int iter;
    for (iter = 0; iter < k_iters && iter < max_iters; ++iter) {
        int i1 = (int)((size_t)orc_rand(&st) % n);
        int i2 = (int)((size_t)orc_rand(&st) % n);
        int i3 = (int)((size_t)orc_rand(&st) % n);
        if (i1 == i2 || i1 == i3 || i2 == i3) continue;
        float p1[3] = {x[i1], y[i1], z[i1]}, p2[3] = {x[i2], y[i2], z[i2]}, p3[3] = {x[i3], y[i3], z[i3]};
        float cand[4];
        if (!orc_plane_model(p1, p2, p3, collinear, cand)) continue;
        int cur = orc_plane_count(x, y, z, n, cand, thr);
        if (cur > best) {
            best = cur;
            memcpy(best_model, cand, sizeof(cand));
            double w = (double)best / (double)n;
            double p_no = 1.0 - pow(w, 3.0);
            if (p_no < DBL_EPSILON) p_no = DBL_EPSILON;
            if (p_no > 1.0 - DBL_EPSILON) p_no = 1.0 - DBL_EPSILON;
            double log_no = log(p_no);
            if (fabs(log_no) > 1e-7) {
                int dyn = (int)ceil(log_p / log_no);
                if (dyn > 0 && dyn < k_iters) k_iters = dyn;
            }
        }
    }
    if (iters_run) *iters_run = iter;
    memcpy(model, best_model, sizeof(best_model));
    return best;
}

/* extract_plane_inliers_outliers_rvv scalar branch (ransac_plane.cpp:374-382). */
void orc_plane_extract(const float *x, const float *y, const float *z, size_t n, const float model[4],
                       float thr, float *inl, size_t *n_in, float *outl, size_t *n_out) {
    float a = model[0], b = model[1], c = model[2], d = model[3];
    size_t ni = 0, no = 0;
    for (size_t j = 0; j < n; ++j) {
        float dist = a * x[j] + b * y[j] + c * z[j] + d;
        if (dist >= -thr && dist <= thr) {
            if (inl) { inl[3 * ni] = x[j]; inl[3 * ni + 1] = y[j]; inl[3 * ni + 2] = z[j]; }
            ni++;
        } else {
            if (outl) { outl[3 * no] = x[j]; outl[3 * no + 1] = y[j]; outl[3 * no + 2] = z[j]; }
            no++;
        }
    }
    *n_in = ni; *n_out = no;
}

/* Euclidean clustering (euclidean_clustering.cpp:113-189, scalar neighbour query :98-109). */
static int cmp_i32(const void *a, const void *b) {
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

size_t orc_euclidean_cluster(const float *x, const float *y, const float *z, size_t n, float tol, int min_size,
                             int max_size, int32_t *indices, int64_t *offsets) {
    offsets[0] = 0;
    if (n == 0) return 0;
    const float tol_sq = tol * tol;
    unsigned char *visited = (unsigned char *)calloc(n, 1);
    int32_t *queue = (int32_t *)malloc(n * sizeof(int32_t));
    size_t ncl = 0, out = 0;
    for (size_t seed = 0; seed < n; ++seed) {
        if (visited[seed]) continue;
        size_t head = 0, tail = 0;
        visited[seed] = 1;
        queue[tail++] = (int32_t)seed;
        while (head < tail) {
            const int32_t cur = queue[head++];
            const float qx = x[cur], qy = y[cur], qz = z[cur];
            for (size_t i = 0; i < n; ++i) {
                if (visited[i]) continue;
                const float dx = x[i] - qx, dy = y[i] - qy, dz = z[i] - qz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (d2 <= tol_sq) {
                    visited[i] = 1;
                    queue[tail++] = (int32_t)i;
                }
            }
        }
        const int sz = (int)tail;
        if (sz >= min_size && sz <= max_size) {
            qsort(queue, tail, sizeof(int32_t), cmp_i32);
            memcpy(indices + out, queue, tail * sizeof(int32_t));
            out += tail;
            offsets[++ncl] = (int64_t)out;
        }
    }
    free(visited);
    free(queue);
    return ncl;
}
