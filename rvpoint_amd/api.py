"""Host-side mirror of the reference's point-cloud API (namespace rvpoint, the "rvv_pcl.h" signatures),
backed by the MI355X kernels through the C ABI.

Names, argument meaning and result conventions follow the reference:
  radius_search_sc / radius_search_rvv        src/search/radius_search.h:11-20
  SpatialHash                                  src/search/spatial_hashing.h:12-53
  sor_sc / sor_rvv / sor_spatial_hash          src/filters/statistical_outlier_removal.h:18-42
  voxel_grid_downsamp_sc / _rvv / _rvv_v2      src/filters/voxel_grid.h:14-34
  normal_estimation_sc / normal_estimation_rvv src/features/normal_estimation.h:15-49
  ransac_plane_sc / ransac_plane_rvv, extract_plane_*   src/segmentation/ransac_plane.h:11-45
Host arrays are numpy; AoS clouds are float32 (n, 3) arrays (PointXYZ), SoA clouds are PointCloudSoA.
Every call stages its inputs into HBM, runs on the GPU and copies results back; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call
from .device import DeviceArray, DeviceCloud


class PointCloudSoA:
    """Non-owning SoA view in the reference (src/core/point_types.h:24-29); here it holds numpy arrays."""

    def __init__(self, x, y, z):
        self.x = np.ascontiguousarray(x, np.float32)
        self.y = np.ascontiguousarray(y, np.float32)
        self.z = np.ascontiguousarray(z, np.float32)
        self.n = len(self.x)

    @classmethod
    def from_aos(cls, aos):
        aos = np.asarray(aos, np.float32).reshape(-1, 3)
        return cls(aos[:, 0], aos[:, 1], aos[:, 2])

    def aos(self):
        return np.stack([self.x, self.y, self.z], 1)


def _soa(cloud):
    if isinstance(cloud, PointCloudSoA):
        return cloud
    return PointCloudSoA.from_aos(cloud)


def _q(query):
    q = np.asarray(query, np.float32).reshape(-1, 3)
    return [np.ascontiguousarray(q[:, i]) for i in range(3)]


# ----------------------------------------------------------------------------------------------------
# H4: brute-force radius search
# ----------------------------------------------------------------------------------------------------
def radius_search_batch_bruteforce(cloud, queries, radius, max_nn):
    """First max_nn hits of every query in index order; returns (idx [nq, max_nn], d2, counts)."""
    c = _soa(cloud)
    qx, qy, qz = _q(queries)
    nq = len(qx)
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    dq = DeviceCloud.from_soa(qx, qy, qz)
    w = max(max_nn, 1)
    idx = DeviceArray((nq, w), np.int32)
    d2 = DeviceArray((nq, w), np.float32)
    cnt = DeviceArray(nq, np.int32)
    call("rvp_radius_bruteforce", *dc.ptrs(), c.n, *dq.ptrs(), nq, float(radius), int(max_nn), idx.ptr, d2.ptr,
         cnt.ptr, None)
    return idx.numpy(), d2.numpy(), cnt.numpy()


def radius_search_sc(cloud_aos, query, radius, max_nn):
    """radius_search_sc (src/search/radius_search.cpp:12-33): (indices, dists) of the first max_nn hits."""
    idx, d2, cnt = radius_search_batch_bruteforce(cloud_aos, [query], radius, max_nn)
    m = int(cnt[0])
    return idx[0, :m].copy(), d2[0, :m].copy()


def radius_search_rvv(cloud_soa, query, radius, max_nn):
    """radius_search_rvv (src/search/radius_search.cpp:38-59); same result as radius_search_sc."""
    if _soa(cloud_soa).n == 0:
        return np.zeros(0, np.int32), np.zeros(0, np.float32)
    return radius_search_sc(cloud_soa, query, radius, max_nn)


# ----------------------------------------------------------------------------------------------------
# H1-H3: SpatialHash
# ----------------------------------------------------------------------------------------------------
class SpatialHash:
    """SpatialHash (src/search/spatial_hashing.h:12-53) with a device-resident grid.

    setInputCloud + build upload the cloud once; radiusSearch answers one query as the reference does,
    radius_search_batch answers many in one launch.  setHashPrimes is accepted for API compatibility; the
    device grid does not hash, so results never depended on the primes."""

    def __init__(self):
        h = C.c_void_p()
        call("rvp_hash_create", C.byref(h))
        self.handle = h.value
        self.cloud = None
        self.dev = None
        self.cell_size = 0.0
        self.primes = (73856093, 19349663)

    def __del__(self):
        try:
            if self.handle:
                from ._lib import lib
                lib().rvp_hash_destroy(self.handle)
        except Exception:
            pass

    def setInputCloud(self, cloud, cell_size):
        self.cloud = _soa(cloud)
        self.cell_size = float(cell_size)

    def setHashPrimes(self, p1, p2):
        self.primes = (int(p1), int(p2))

    def build(self):
        c = self.cloud
        self.dev = DeviceCloud.from_soa(c.x, c.y, c.z)
        call("rvp_hash_build", self.handle, *self.dev.ptrs(), c.n, self.cell_size, None)

    @classmethod
    def from_device(cls, dev: DeviceCloud, cell_size, stream=None):
        """Build on a cloud already resident in HBM."""
        h = cls()
        h.dev = dev
        h.cell_size = float(cell_size)
        call("rvp_hash_build", h.handle, *dev.ptrs(), dev.n, h.cell_size, stream)
        return h

    # -- introspection
    def bbox(self):
        out = (C.c_float * 6)()
        call("rvp_hash_bbox", self.handle, out)
        return np.array(out, np.float32)

    def info(self):
        n, occ = C.c_size_t(), C.c_size_t()
        dims = (C.c_int * 3)()
        cell = C.c_float()
        call("rvp_hash_info", self.handle, C.byref(n), C.byref(occ), dims, C.byref(cell))
        return {"n": n.value, "occupied": occ.value, "dims": tuple(dims), "grid_cell": cell.value}

    def cell_coords(self):
        n = self.dev.n
        ixyz = DeviceArray((n, 3), np.int32)
        keys = DeviceArray(n, np.int64)
        call("rvp_hash_cell_coords", self.handle, ixyz.ptr, keys.ptr, None)
        return ixyz.numpy(), keys.numpy()

    # -- queries
    def radius_search_batch(self, queries, radius, max_nn=0, stream=None):
        """Batched radiusSearch: list of (indices, dists) per query, each sorted by (d2, idx), truncated to
        max_nn when max_nn > 0.  queries=None queries the cloud's own points."""
        if queries is None:
            nq, qp = self.dev.n, (None, None, None)
            dq = None
        else:
            qx, qy, qz = _q(queries)
            nq = len(qx)
            dq = DeviceCloud.from_soa(qx, qy, qz)
            qp = dq.ptrs()
        if nq == 0:
            return []
        if max_nn > 0:
            idx = DeviceArray((nq, max_nn), np.int32)
            d2 = DeviceArray((nq, max_nn), np.float32)
            cnt = DeviceArray(nq, np.int32)
            call("rvp_radius_search", self.handle, *qp, nq, float(radius), int(max_nn), idx.ptr, d2.ptr, cnt.ptr,
                 None, stream)
            I, D, N = idx.numpy(stream=stream), d2.numpy(stream=stream), cnt.numpy(stream=stream)
            return [(I[i, :N[i]].copy(), D[i, :N[i]].copy()) for i in range(nq)]
        cnt = DeviceArray(nq, np.int32)
        call("rvp_radius_count", self.handle, *qp, nq, float(radius), cnt.ptr, stream)
        off = DeviceArray(nq + 1, np.int64)
        tot = C.c_int64()
        call("rvp_exclusive_scan", cnt.ptr, nq, off.ptr, C.byref(tot), stream)
        idx = DeviceArray(max(tot.value, 1), np.int32)
        d2 = DeviceArray(max(tot.value, 1), np.float32)
        call("rvp_radius_fill", self.handle, *qp, nq, float(radius), off.ptr, idx.ptr, d2.ptr, stream)
        O = off.numpy(stream=stream)
        I, D = idx.numpy(stream=stream), d2.numpy(stream=stream)
        return [(I[O[i]:O[i + 1]].copy(), D[O[i]:O[i + 1]].copy()) for i in range(nq)]

    def radiusSearch(self, query, radius, max_nn=0):
        """SpatialHash::radiusSearch (spatial_hashing.cpp:165-223) for one query: (indices, dists)."""
        return self.radius_search_batch([query], radius, max_nn)[0]

    def knn(self, queries, k):
        """k nearest by (d2, idx) of each query (rows padded with -1 / inf when the cloud is smaller)."""
        if queries is None:
            nq, qp = self.dev.n, (None, None, None)
        else:
            qx, qy, qz = _q(queries)
            nq = len(qx)
            dq = DeviceCloud.from_soa(qx, qy, qz)
            qp = dq.ptrs()
        idx = DeviceArray((nq, k), np.int32)
        d2 = DeviceArray((nq, k), np.float32)
        call("rvp_knn", self.handle, *qp, nq, int(k), idx.ptr, d2.ptr, None)
        return idx.numpy(), d2.numpy()


# ----------------------------------------------------------------------------------------------------
# SOR
# ----------------------------------------------------------------------------------------------------
def _sor_exact(c: PointCloudSoA, k, alpha, with_means=False):
    n = c.n
    if n == 0:
        return (np.zeros((0, 3), np.float32), np.zeros(0, np.float32)) if with_means else np.zeros((0, 3), np.float32)
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    out = DeviceArray((n, 3), np.float32)
    means = DeviceArray(n, np.float32)
    m = C.c_size_t()
    call("rvp_sor_exact", *dc.ptrs(), n, int(k), float(alpha), out.ptr, C.byref(m), None, means.ptr, None)
    res = out.numpy()[:m.value]
    return (res, means.numpy()) if with_means else res


def sor_sc(cloud_aos, k, alpha, with_means=False):
    """sor_sc (statistical_outlier_removal.cpp:19-63): kept points (AoS) in input order."""
    return _sor_exact(_soa(cloud_aos), k, alpha, with_means)


def sor_rvv(cloud_soa, k, alpha, with_means=False):
    """sor_rvv (statistical_outlier_removal.cpp:68-190); same result as sor_sc."""
    return _sor_exact(_soa(cloud_soa), k, alpha, with_means)


def sor_spatial_hash(cloud_soa, hash_: SpatialHash, k, alpha, search_radius=0.5, with_means=False):
    """sor_spatial_hash (statistical_outlier_removal.cpp:246-270); `hash_` is built on the same cloud."""
    n = hash_.dev.n
    out = DeviceArray((max(n, 1), 3), np.float32)
    means = DeviceArray(max(n, 1), np.float32)
    m = C.c_size_t()
    call("rvp_sor_radius", hash_.handle, int(k), float(alpha), float(search_radius), out.ptr, C.byref(m), None,
         means.ptr, None)
    res = out.numpy()[:m.value]
    return (res, means.numpy()[:n]) if with_means else res


# ----------------------------------------------------------------------------------------------------
# voxel grid
# ----------------------------------------------------------------------------------------------------
def _voxel(c: PointCloudSoA, leaf):
    if c.n == 0:
        return np.zeros((0, 3), np.float32)
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    out = DeviceArray((c.n, 3), np.float32)
    m = C.c_size_t()
    call("rvp_voxel", *dc.ptrs(), c.n, float(leaf), out.ptr, C.byref(m), None)
    return out.numpy()[:m.value]


def voxel_grid_downsamp_sc(cloud_aos, leaf_size):
    """voxel_grid_downsamp_sc (voxel_grid_downsamp.cpp:183-213): centroids in (vx, vy, vz) order."""
    return _voxel(_soa(cloud_aos), leaf_size)


def voxel_grid_downsamp_rvv(cloud_soa, leaf_size):
    return _voxel(_soa(cloud_soa), leaf_size)


def voxel_grid_downsamp_rvv_v2(cloud_soa, leaf_size):
    return _voxel(_soa(cloud_soa), leaf_size)


# ----------------------------------------------------------------------------------------------------
# normals
# ----------------------------------------------------------------------------------------------------
def _normals_out(n):
    return [DeviceArray(max(n, 1), np.float32) for _ in range(3)]


def normal_estimation_sc(cloud_aos, k, radius, vp_x=0.0, vp_y=0.0, vp_z=0.0, eigen_iters=4):
    """normal_estimation_sc (normal_estimation.cpp:72-112): (nx, ny, nz).  `radius` and `eigen_iters` are
    accepted for signature compatibility (the reference ignores radius here; eigen solves converge)."""
    c = _soa(cloud_aos)
    if c.n == 0:
        return tuple(np.zeros(0, np.float32) for _ in range(3))
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    nx, ny, nz = _normals_out(c.n)
    call("rvp_normals_knn", *dc.ptrs(), c.n, int(k), float(vp_x), float(vp_y), float(vp_z), nx.ptr, ny.ptr, nz.ptr,
         None)
    return nx.numpy()[:c.n], ny.numpy()[:c.n], nz.numpy()[:c.n]


def normal_estimation_rvv(cloud_soa, hash_: SpatialHash, k, radius, vp_x=0.0, vp_y=0.0, vp_z=0.0, eigen_iters=4,
                          max_nn=0):
    """normal_estimation_rvv(in, SpatialHash&, ...) (normal_estimation.cpp:189-214): all neighbours within
    radius (max_nn > 0 caps the (d2, idx)-sorted list); fewer than 3 gives a zero normal."""
    n = hash_.dev.n
    if n == 0:
        return tuple(np.zeros(0, np.float32) for _ in range(3))
    nx, ny, nz = _normals_out(n)
    call("rvp_normals_radius", hash_.handle, float(radius), int(max_nn), float(vp_x), float(vp_y), float(vp_z),
         nx.ptr, ny.ptr, nz.ptr, None)
    return nx.numpy()[:n], ny.numpy()[:n], nz.numpy()[:n]


def normals_fallback_count():
    c = C.c_size_t()
    call("rvp_normals_fallback_count", C.byref(c))
    return c.value


# ----------------------------------------------------------------------------------------------------
# RANSAC + extraction
# ----------------------------------------------------------------------------------------------------
def _ransac(c: PointCloudSoA, dist_thresh, max_iters, collinear_thresh, probability):
    model = (C.c_float * 4)()
    best = C.c_int()
    iters = C.c_int()
    if c.n < 3:
        return 0, np.zeros(4, np.float32), 0
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    call("rvp_ransac_plane", *dc.ptrs(), c.n, float(dist_thresh), int(max_iters), float(collinear_thresh),
         float(probability), model, C.byref(best), C.byref(iters), None)
    return best.value, np.array(model, np.float32), iters.value


def ransac_plane_sc(cloud_aos, dist_thresh, max_iters, collinear_thresh=1e-6, probability=0.99):
    """ransac_plane_sc (ransac_plane.cpp:55-111): (inliers, model[4])."""
    best, model, _ = _ransac(_soa(cloud_aos), dist_thresh, max_iters, collinear_thresh, probability)
    return best, model


def ransac_plane_rvv(cloud_soa, dist_thresh, max_iters, collinear_thresh=1e-6, probability=0.99):
    best, model, _ = _ransac(_soa(cloud_soa), dist_thresh, max_iters, collinear_thresh, probability)
    return best, model


def ransac_plane_extract(cloud_soa, dist_thresh, max_iters, collinear_thresh=1e-6, probability=0.99):
    """RANSAC then the inlier / outlier split by the best model in one call, the stage split of
    pipeline_export.cpp:457-498: (inliers, model[4], iterations, inliers AoS, outliers AoS).  The split runs
    only when a model was found."""
    c = _soa(cloud_soa)
    empty = np.zeros((0, 3), np.float32)
    if c.n < 3:
        return 0, np.zeros(4, np.float32), 0, empty, empty
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    a = DeviceArray((c.n, 3), np.float32)
    b = DeviceArray((c.n, 3), np.float32)
    model = (C.c_float * 4)()
    best, iters = C.c_int(), C.c_int()
    ni, no = C.c_size_t(), C.c_size_t()
    call("rvp_ransac_plane_extract", *dc.ptrs(), c.n, float(dist_thresh), int(max_iters), float(collinear_thresh),
         float(probability), model, C.byref(best), C.byref(iters), a.ptr, C.byref(ni), b.ptr, C.byref(no), None)
    return best.value, np.array(model, np.float32), iters.value, a.numpy()[:ni.value], b.numpy()[:no.value]


def extract_plane_inliers_outliers_rvv(cloud_soa, model, dist_thresh):
    """extract_plane_inliers_outliers_rvv (ransac_plane.cpp:374-382): (inliers AoS, outliers AoS)."""
    c = _soa(cloud_soa)
    if c.n == 0:
        z = np.zeros((0, 3), np.float32)
        return z, z
    dc = DeviceCloud.from_soa(c.x, c.y, c.z)
    a = DeviceArray((c.n, 3), np.float32)
    b = DeviceArray((c.n, 3), np.float32)
    m = (C.c_float * 4)(*[float(v) for v in np.asarray(model, np.float32)])
    ni, no = C.c_size_t(), C.c_size_t()
    call("rvp_plane_extract", *dc.ptrs(), c.n, m, float(dist_thresh), a.ptr, C.byref(ni), b.ptr, C.byref(no), None)
    return a.numpy()[:ni.value], b.numpy()[:no.value]


def extract_plane_inliers_rvv(cloud_soa, model, dist_thresh):
    return extract_plane_inliers_outliers_rvv(cloud_soa, model, dist_thresh)[0]


def extract_plane_outliers_rvv(cloud_soa, model, dist_thresh):
    return extract_plane_inliers_outliers_rvv(cloud_soa, model, dist_thresh)[1]


# ----------------------------------------------------------------------------------------------------
# batches: many clouds through one launch per stage
# ----------------------------------------------------------------------------------------------------
class Batch:
    """Clouds back to back in HBM with device offsets (the batched workload: one launch per stage)."""

    def __init__(self, clouds, stream=None):
        clouds = [np.asarray(c, np.float32).reshape(-1, 3) for c in clouds]
        self.sizes = np.array([len(c) for c in clouds], np.int64)
        self.offsets_host = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.uint32)
        allp = np.concatenate(clouds) if clouds else np.zeros((0, 3), np.float32)
        self.n = len(allp)
        self.ncl = len(clouds)
        self.dev = DeviceCloud(allp if self.n else np.zeros((1, 3), np.float32), stream)
        self.offsets = DeviceArray.from_numpy(self.offsets_host, stream)

    def split(self, arr):
        return [arr[self.offsets_host[c]:self.offsets_host[c + 1]] for c in range(self.ncl)]


def batch_sor_exact(batch: Batch, k, alpha, stream=None):
    """S1 per cloud in one pass: (list of kept AoS arrays, means per point)."""
    nb = max(batch.n, 1)
    out = DeviceArray((nb, 3), np.float32)
    out_off = DeviceArray(batch.ncl + 1, np.uint32)
    means = DeviceArray(nb, np.float32)
    call("rvp_batch_sor_exact", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, nb, int(k), float(alpha), out.ptr,
         out_off.ptr, None, means.ptr, stream)
    o = out_off.numpy(stream=stream)
    res = out.numpy(stream=stream)
    return [res[o[c]:o[c + 1]].copy() for c in range(batch.ncl)], means.numpy(stream=stream)[:batch.n]


def batch_radius_search(batch: Batch, radius, max_nn, cell=None, stream=None):
    """H3 per cloud: (idx rows [n, max_nn] with cloud-local indices, d2 rows, counts)."""
    nb = max(batch.n, 1)
    r = np.broadcast_to(np.asarray(radius, np.float32), (batch.ncl,)).copy()
    dr = DeviceArray.from_numpy(r, stream)
    dc = DeviceArray.from_numpy(np.broadcast_to(np.asarray(cell, np.float32), (batch.ncl,)).copy(), stream) \
        if cell is not None else None
    idx = DeviceArray((nb, max_nn), np.int32)
    d2 = DeviceArray((nb, max_nn), np.float32)
    cnt = DeviceArray(nb, np.int32)
    call("rvp_batch_radius_search", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, nb, dr.ptr,
         dc.ptr if dc is not None else None, int(max_nn), idx.ptr, d2.ptr, cnt.ptr, stream)
    return tuple(a.numpy(stream=stream)[:batch.n] for a in (idx, d2, cnt))


def batch_normals_knn(batch: Batch, k, vp=(0.0, 0.0, 0.0), stream=None):
    """N1 per cloud: (n, 3) normals by batch position."""
    nb = max(batch.n, 1)
    nx, ny, nz = (DeviceArray(nb, np.float32) for _ in range(3))
    call("rvp_batch_normals_knn", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, nb, int(k), float(vp[0]),
         float(vp[1]), float(vp[2]), nx.ptr, ny.ptr, nz.ptr, stream)
    return np.stack([a.numpy(stream=stream)[:batch.n] for a in (nx, ny, nz)], 1)


def batch_normals_knn_lists(batch: Batch, k, idx, d2, cnt, vp=(0.0, 0.0, 0.0), stream=None):
    """N1 per cloud as batch_normals_knn, each query's k + 1 nearest read from its capped radius list of the same
    clouds (idx (n, cap) and d2 (n, cap) or None, cnt (n,), as batch_radius_search returns them) when the list
    holds at least k + 1 entries: (n, 3) normals by batch position."""
    nb = max(batch.n, 1)
    idx = np.ascontiguousarray(idx, np.int32)
    cap = int(idx.shape[1])
    pad = lambda a, dt: np.concatenate([np.asarray(a, dt), np.zeros((nb - len(a),) + np.shape(a)[1:], dt)])
    di = DeviceArray.from_numpy(pad(idx, np.int32), stream)
    dd = DeviceArray.from_numpy(pad(d2, np.float32), stream) if d2 is not None else None
    dc = DeviceArray.from_numpy(pad(cnt, np.int32), stream)
    nx, ny, nz = (DeviceArray(nb, np.float32) for _ in range(3))
    call("rvp_batch_normals_knn_lists", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, nb, int(k), di.ptr,
         dd.ptr if dd is not None else None, dc.ptr, cap, float(vp[0]), float(vp[1]), float(vp[2]), nx.ptr, ny.ptr,
         nz.ptr, stream)
    return np.stack([a.numpy(stream=stream)[:batch.n] for a in (nx, ny, nz)], 1)


def batch_normals_radius(batch: Batch, radius, max_nn, vp=(0.0, 0.0, 0.0), stream=None, cell=None):
    """N2 per cloud with a cap (radius per cloud or scalar): (n, 3) normals by batch position.  cell (per cloud or
    scalar, optional): the grid's cell edge instead of the radius."""
    nb = max(batch.n, 1)
    r = DeviceArray.from_numpy(np.broadcast_to(np.asarray(radius, np.float32), (batch.ncl,)).copy(), stream)
    ce = None if cell is None else DeviceArray.from_numpy(
        np.broadcast_to(np.asarray(cell, np.float32), (batch.ncl,)).copy(), stream)
    nx, ny, nz = (DeviceArray(nb, np.float32) for _ in range(3))
    call("rvp_batch_normals_radius", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, nb, r.ptr,
         None if ce is None else ce.ptr, int(max_nn), float(vp[0]), float(vp[1]), float(vp[2]), nx.ptr, ny.ptr,
         nz.ptr, stream)
    return np.stack([a.numpy(stream=stream)[:batch.n] for a in (nx, ny, nz)], 1)


def batch_voxel(batch: Batch, leaf, stream=None):
    """V1 per cloud (leaf per cloud or scalar): (list of AoS voxel centroids, status per cloud)."""
    nb = max(batch.n, 1)
    lf = DeviceArray.from_numpy(np.broadcast_to(np.asarray(leaf, np.float32), (batch.ncl,)).copy(), stream)
    out = DeviceArray((nb, 3), np.float32)
    out_off = DeviceArray(batch.ncl + 1, np.uint32)
    status = DeviceArray(batch.ncl, np.uint32)
    call("rvp_batch_voxel", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, nb, lf.ptr, out.ptr, out_off.ptr,
         status.ptr, stream)
    o = out_off.numpy(stream=stream)
    res = out.numpy(stream=stream)
    return [res[o[c]:o[c + 1]].copy() for c in range(batch.ncl)], status.numpy(stream=stream)


def batch_ransac_plane(batch: Batch, dist_thresh, max_iters, collinear_thresh=1e-6, probability=0.99, stream=None):
    """R1 per cloud: (models [ncl, 4], inliers [ncl], iterations [ncl])."""
    m = np.zeros((batch.ncl, 4), np.float32)
    inl = np.zeros(batch.ncl, np.int32)
    its = np.zeros(batch.ncl, np.int32)
    call("rvp_batch_ransac_plane", *batch.dev.ptrs(), batch.offsets.ptr, batch.ncl, max(batch.n, 1),
         float(dist_thresh), int(max_iters), float(collinear_thresh), float(probability),
         m.ctypes.data_as(C.POINTER(C.c_float)), inl.ctypes.data_as(C.POINTER(C.c_int)),
         its.ctypes.data_as(C.POINTER(C.c_int)), stream)
    return m, inl, its


# ----------------------------------------------------------------------------------------------------
# Euclidean clustering
# ----------------------------------------------------------------------------------------------------
def euclidean_cluster(cloud, tolerance, min_size=1, max_size=2**31 - 1, hash_=None, stream=None):
    """EuclideanClustering::extract (euclidean_clustering.cpp:113-189): sorted index arrays, one per cluster of
    min_size..max_size points, in the order of their smallest index.  hash_: a SpatialHash already built on
    the cloud (any cell); else one with cell = tolerance is built."""
    if hash_ is None:
        c = _soa(cloud)
        if c.n == 0:
            return []
        hash_ = SpatialHash()
        ext = float(max(np.ptp(c.x), np.ptp(c.y), np.ptp(c.z), 1e-30))
        hash_.setInputCloud(c, max(float(tolerance), ext * 1e-6))
        hash_.build()
    n = hash_.dev.n
    if n == 0:
        return []
    idx = DeviceArray(n, np.int32)
    off = DeviceArray(n + 1, np.int64)
    m = C.c_size_t()
    call("rvp_euclidean_cluster", hash_.handle, float(tolerance), int(min_size), int(max_size), idx.ptr, off.ptr,
         C.byref(m), None, stream)
    o = off.numpy(stream=stream)[:m.value + 1]
    I = idx.numpy(stream=stream)
    return [I[o[i]:o[i + 1]].copy() for i in range(m.value)]


# ----------------------------------------------------------------------------------------------------
# PCD ingest
# ----------------------------------------------------------------------------------------------------
def pcd_info(path):
    n = C.c_size_t()
    mode = C.c_int()
    call("rvp_pcd_info", str(path).encode(), C.byref(n), C.byref(mode))
    return n.value, {0: "ascii", 1: "binary", 2: "binary_compressed"}.get(mode.value, "other")


def load_pcd(path):
    """loadPCD (simple_pcd_loader.h:143-293): (n, 3) float32 xyz."""
    n, _ = pcd_info(path)
    x, y, z = (np.zeros(max(n, 1), np.float32) for _ in range(3))
    m = C.c_size_t()
    P = C.POINTER(C.c_float)
    call("rvp_pcd_read", str(path).encode(), x.ctypes.data_as(P), y.ctypes.data_as(P), z.ctypes.data_as(P), n,
         C.byref(m))
    k = m.value
    return np.stack([x[:k], y[:k], z[:k]], 1)


def load_pcd_device(path, stream=None):
    """The same straight into device SoA (a DeviceCloud)."""
    n, _ = pcd_info(path)
    d = DeviceCloud.__new__(DeviceCloud)
    d.x, d.y, d.z = (DeviceArray(max(n, 1), np.float32) for _ in range(3))
    m = C.c_size_t()
    call("rvp_pcd_read_device", str(path).encode(), d.x.ptr, d.y.ptr, d.z.ptr, n, C.byref(m), stream)
    d.n = m.value
    return d


def save_pcd(path, xyz, binary=False):
    """savePCD (simple_pcd_loader.h:295-328): FIELDS x y z, ascii or binary."""
    c = _soa(xyz)
    P = C.POINTER(C.c_float)
    call("rvp_pcd_write", str(path).encode(), c.x.ctypes.data_as(P), c.y.ctypes.data_as(P), c.z.ctypes.data_as(P),
         c.n, 1 if binary else 0)
