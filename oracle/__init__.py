"""ctypes front-end of the CPU oracle (oracle.c) and of oracle/_ref (the reference's own scalar path).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by bench.py's cpu_baseline leg,
always as the checker or as the timed CPU baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "librvpoint_ref.so")
REF_SRC = "/root/reference"

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_sz = C.c_size_t


def build(with_ref: bool | None = None) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir(REF_SRC)
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REF_SRC}"], check=True)
        if os.path.exists(os.path.join(os.path.dirname(HERE), "rvpoint_amd", "lib", "librvpoint_amd.so")):
            # the reference's unit tests against the drop-in library (run by the GPU tests)
            subprocess.run(["make", "-s", "-C", HERE, "reftests", f"REF={REF_SRC}"], check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(with_ref=False)
        L = C.CDLL(LIB_PATH)
        L.orc_dist_sq.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_float, C.c_float, _f32p]
        L.orc_inds_in_radius.argtypes = [_f32p, _f32p, _f32p, _i32p, _sz, C.c_float, C.c_float, C.c_float, C.c_float,
                                         _i32p, _f32p, _sz]
        L.orc_inds_in_radius.restype = _sz
        L.orc_bbox.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, _f32p]
        L.orc_cell_coords.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, _i32p, C.c_void_p]
        L.orc_hash_radius.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, _f32p, _f32p, _f32p, _sz,
                                      C.c_float, C.c_int, _sz, C.c_void_p, C.c_void_p, _i32p, C.c_void_p]
        L.orc_radius_all.argtypes = L.orc_hash_radius.argtypes[:4] + L.orc_hash_radius.argtypes[5:]
        L.orc_radius_bruteforce.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_float, C.c_float,
                                            C.c_float, _i32p, _f32p, C.c_int]
        L.orc_radius_bruteforce.restype = _sz
        L.orc_knn.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, _f32p, _f32p, _sz, C.c_int, _i32p, _f32p]
        L.orc_sor_sc.argtypes = [_f32p, _sz, C.c_int, C.c_float, _f32p, _f32p]
        L.orc_sor_sc.restype = _sz
        L.orc_sor_sc_mt.argtypes = [_f32p, _sz, C.c_int, C.c_float, _f32p, _f32p, C.c_int]
        L.orc_sor_sc_mt.restype = _sz
        L.orc_normals_hash_e1_mt.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_float, _f32p, C.c_int,
                                             _f32p, _f32p, _f32p, C.c_int]
        L.orc_sor_radius.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_int, C.c_float, C.c_float,
                                     _f32p, _f32p]
        L.orc_sor_radius.restype = _sz
        L.orc_sor_filter.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_float, _f32p, _u8p, _f32p]
        L.orc_sor_filter.restype = _sz
        L.orc_seq_sum.argtypes = [_f32p, _sz, C.c_int, C.c_float]
        L.orc_seq_sum.restype = C.c_float
        L.orc_voxel_sc.argtypes = [_f32p, _sz, C.c_float, _f32p]
        L.orc_voxel_sc.restype = _sz
        L.orc_normals_knn_f64.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_int, _f32p, _f64p, _f64p, _f64p]
        L.orc_normals_radius_f64.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_float, C.c_int, _f32p,
                                             _f64p, _f64p, _f64p]
        L.orc_normal_from_points_f64.argtypes = [_f64p, _f64p, _f64p, _sz, _f64p, _f64p]
        L.orc_eigen_e1.argtypes = [_f32p, C.c_int, _f32p]
        L.orc_flip.argtypes = [C.c_float] * 6 + [_f32p]
        L.orc_srand.argtypes = [C.c_void_p, C.c_uint]
        L.orc_rand.argtypes = [C.c_void_p]
        L.orc_rand.restype = C.c_int32
        L.orc_plane_model.argtypes = [_f32p, _f32p, _f32p, C.c_float, _f32p]
        L.orc_plane_count.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_float]
        L.orc_ransac_plane.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_int, C.c_float, C.c_float,
                                       _f32p, C.POINTER(C.c_int)]
        L.orc_plane_extract.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_float, _f32p,
                                        C.POINTER(_sz), _f32p, C.POINTER(_sz)]
        L.orc_euclidean_cluster.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_int, C.c_int,
                                            np.ctypeslib.ndpointer(np.int32, flags="C"),
                                            np.ctypeslib.ndpointer(np.int64, flags="C")]
        L.orc_euclidean_cluster.restype = _sz
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own scalar path (oracle/_ref). Raises when it was not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_PATH)
        R = C.CDLL(REF_PATH)
        R.ref_get_dist_sq.argtypes = [_f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float, _f32p, _sz]
        R.ref_inds_in_radius.argtypes = [_f32p, _f32p, _f32p, _i32p, _sz, C.c_float, C.c_float, C.c_float,
                                         C.c_float, _i32p, _f32p, _sz]
        R.ref_inds_in_radius.restype = _sz
        R.ref_hash_radius.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, _f32p, _f32p, _f32p, _sz, C.c_float,
                                      C.c_int, _sz, C.c_void_p, C.c_void_p, _i32p]
        R.ref_radius_search_sc.argtypes = [_f32p, _sz, C.c_float, C.c_float, C.c_float, C.c_float, _i32p, _f32p,
                                           C.c_int]
        R.ref_radius_search_sc.restype = _sz
        R.ref_radius_search_rvv.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_float, C.c_float,
                                            C.c_float, _i32p, _f32p, C.c_int]
        R.ref_radius_search_rvv.restype = _sz
        R.ref_sor_sc.argtypes = [_f32p, _sz, _f32p, C.c_int, C.c_float]
        R.ref_sor_sc.restype = _sz
        R.ref_sor_rvv.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_int, C.c_float]
        R.ref_sor_rvv.restype = _sz
        R.ref_sor_spatial_hash.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, _f32p, C.c_int, C.c_float,
                                           C.c_float]
        R.ref_sor_spatial_hash.restype = _sz
        R.ref_sor_octree.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_int, C.c_float, C.c_float, C.c_int]
        R.ref_sor_octree.restype = _sz
        R.ref_octree_radius.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_float, C.c_float, C.c_float,
                                        C.c_int, np.ctypeslib.ndpointer(np.int32, flags="C"), _f32p, _sz]
        R.ref_octree_radius.restype = _sz
        R.ref_normal_octree.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, _f32p, _f32p, C.c_int, C.c_float,
                                        C.c_float, C.c_float, C.c_float, C.c_int]
        R.ref_voxel_sc.argtypes = [_f32p, _sz, _f32p, C.c_float]
        R.ref_voxel_sc.restype = _sz
        R.ref_voxel_v2.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_float]
        R.ref_voxel_v2.restype = _sz
        R.ref_normal_sc.argtypes = [_f32p, _sz, _f32p, _f32p, _f32p, C.c_int, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int]
        R.ref_normal_hash.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, _f32p, _f32p, _f32p, C.c_int,
                                      C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]
        R.ref_ransac_sc.argtypes = [_f32p, _sz, C.c_float, C.c_int, _f32p, C.c_float, C.c_float]
        R.ref_ransac_rvv.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_int, _f32p, C.c_float, C.c_float]
        R.ref_extract.argtypes = [_f32p, _f32p, _f32p, _sz, _f32p, C.c_float, _f32p, _f32p, C.POINTER(_sz),
                                  C.POINTER(_sz)]
        R.ref_euclidean_cluster.argtypes = [_f32p, _f32p, _f32p, _sz, C.c_float, C.c_int, C.c_int,
                                            np.ctypeslib.ndpointer(np.int32, flags="C"),
                                            np.ctypeslib.ndpointer(np.int64, flags="C")]
        R.ref_euclidean_cluster.restype = _sz
        R.ref_load_pcd.argtypes = [C.c_char_p, _f32p, _sz]
        R.ref_load_pcd.restype = C.c_int
        R.ref_save_pcd.argtypes = [C.c_char_p, _f32p, _sz, C.c_int]
        _ref = R
    return _ref


# ---------------------------------------------------------------------------------------------------------
# numpy-level helpers (oracle)
# ---------------------------------------------------------------------------------------------------------

def _c(a, dt=np.float32):
    return np.ascontiguousarray(a, dtype=dt)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dist_sq(x, y, z, q):
    out = np.zeros(len(x), np.float32)
    lib().orc_dist_sq(_c(x), _c(y), _c(z), len(x), q[0], q[1], q[2], out)
    return out


def inds_in_radius(x, y, z, subset, q, r2, prefix_idx=(), prefix_d2=()):
    """D2 appending after the given prefix; returns (indices, d2) including the prefix."""
    subset = _c(subset, np.int32)
    have = len(prefix_idx)
    oi = np.zeros(have + len(subset), np.int32)
    od = np.zeros(have + len(subset), np.float32)
    oi[:have] = prefix_idx
    od[:have] = prefix_d2
    m = lib().orc_inds_in_radius(_c(x), _c(y), _c(z), subset, len(subset), q[0], q[1], q[2], r2, oi, od, have)
    return oi[:m], od[:m]


def bbox(x, y, z, cell):
    out = np.zeros(6, np.float32)
    lib().orc_bbox(_c(x), _c(y), _c(z), len(x), cell, out)
    return out


def cell_coords(x, y, z, cell):
    n = len(x)
    ixyz = np.zeros((n, 3), np.int32)
    keys = np.zeros(n, np.int64)
    lib().orc_cell_coords(_c(x), _c(y), _c(z), n, cell, ixyz.reshape(-1), _ptr(keys))
    return ixyz, keys


def _rows(fn, x, y, z, extra, qx, qy, qz, radius, max_nn):
    nq = len(qx)
    counts = np.zeros(nq, np.int32)
    total = np.zeros(nq, np.int32)
    args = [_c(x), _c(y), _c(z), len(x)] + extra + [_c(qx), _c(qy), _c(qz), nq, radius, max_nn]
    fn(*args, 0, None, None, counts, _ptr(total))
    stride = int(max(1, counts.max() if nq else 1))
    idx = np.full((nq, stride), -1, np.int32)
    d2 = np.full((nq, stride), np.inf, np.float32)
    fn(*args, stride, _ptr(idx), _ptr(d2), counts, _ptr(total))
    return idx, d2, counts, total


def hash_radius(x, y, z, cell, qx, qy, qz, radius, max_nn=0):
    """H3 lists (idx, d2 rows padded with -1/inf, counts, untruncated totals)."""
    return _rows(lib().orc_hash_radius, x, y, z, [C.c_float(cell)], qx, qy, qz, radius, max_nn)


def radius_all(x, y, z, qx, qy, qz, radius, max_nn=0):
    return _rows(lib().orc_radius_all, x, y, z, [], qx, qy, qz, radius, max_nn)


def radius_bruteforce(x, y, z, q, radius, max_nn):
    idx = np.zeros(max(max_nn, 1), np.int32)
    d = np.zeros(max(max_nn, 1), np.float32)
    m = lib().orc_radius_bruteforce(_c(x), _c(y), _c(z), len(x), q[0], q[1], q[2], radius, idx, d, max_nn)
    return idx[:m], d[:m]


def knn(x, y, z, qx, qy, qz, k):
    nq = len(qx)
    idx = np.zeros((nq, k), np.int32)
    d2 = np.zeros((nq, k), np.float32)
    lib().orc_knn(_c(x), _c(y), _c(z), len(x), _c(qx), _c(qy), _c(qz), nq, k, idx.reshape(-1), d2.reshape(-1))
    return idx, d2


def sor_sc(aos, k, alpha):
    aos = _c(aos).reshape(-1, 3)
    n = len(aos)
    out = np.zeros((max(n, 1), 3), np.float32)
    means = np.zeros(max(n, 1), np.float32)
    m = lib().orc_sor_sc(aos.reshape(-1), n, k, alpha, out.reshape(-1), means)
    return out[:m], means[:n]


def threads() -> int:
    """Host threads the multi-threaded oracle functions use: the CPUs this process may run on."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except (AttributeError, OSError):
        return max(1, os.cpu_count() or 1)


def sor_sc_mt(aos, k, alpha, nthreads=None):
    """sor_sc with the per-point means on several threads; bit-identical to sor_sc."""
    aos = _c(aos).reshape(-1, 3)
    n = len(aos)
    out = np.zeros((max(n, 1), 3), np.float32)
    means = np.zeros(max(n, 1), np.float32)
    m = lib().orc_sor_sc_mt(aos.reshape(-1), n, k, alpha, out.reshape(-1), means, nthreads or threads())
    return out[:m], means[:n]


def normals_hash_e1(x, y, z, cell, radius, vp=(0.0, 0.0, 0.0), eigen_iters=4, nthreads=1):
    """normal_estimation_rvv with a SpatialHash, reference semantics (uncapped radius list, E1, flip)."""
    n = len(x)
    nx, ny, nz = (np.zeros(n, np.float32) for _ in range(3))
    lib().orc_normals_hash_e1_mt(_c(x), _c(y), _c(z), n, cell, radius, _c(vp), eigen_iters, nx, ny, nz, nthreads)
    return nx, ny, nz


def sor_radius(x, y, z, cell, k, alpha, search_radius):
    n = len(x)
    out = np.zeros((max(n, 1), 3), np.float32)
    means = np.zeros(max(n, 1), np.float32)
    m = lib().orc_sor_radius(_c(x), _c(y), _c(z), n, cell, k, alpha, search_radius, out.reshape(-1), means)
    return out[:m], means[:n]


def sor_filter(x, y, z, means, alpha):
    n = len(x)
    out = np.zeros((max(n, 1), 3), np.float32)
    keep = np.zeros(max(n, 1), np.uint8)
    stats = np.zeros(3, np.float32)
    m = lib().orc_sor_filter(_c(x), _c(y), _c(z), n, _c(means), alpha, out.reshape(-1), keep, stats)
    return out[:m], keep[:n].astype(bool), stats


def seq_sum(v, squared_dev=False, mean=0.0):
    v = _c(v)
    return np.float32(lib().orc_seq_sum(v, len(v), 1 if squared_dev else 0, mean))


def voxel_sc(aos, leaf):
    aos = _c(aos).reshape(-1, 3)
    out = np.zeros((max(len(aos), 1), 3), np.float32)
    m = lib().orc_voxel_sc(aos.reshape(-1), len(aos), leaf, out.reshape(-1))
    return out[:m]


def normals_knn_f64(x, y, z, k, vp=(0.0, 0.0, 0.0)):
    n = len(x)
    nx, ny, nz = (np.zeros(n, np.float64) for _ in range(3))
    lib().orc_normals_knn_f64(_c(x), _c(y), _c(z), n, k, _c(vp), nx, ny, nz)
    return np.stack([nx, ny, nz], 1)


def normals_radius_f64(x, y, z, cell, radius, max_nn=0, vp=(0.0, 0.0, 0.0)):
    n = len(x)
    nx, ny, nz = (np.zeros(n, np.float64) for _ in range(3))
    lib().orc_normals_radius_f64(_c(x), _c(y), _c(z), n, cell, radius, max_nn, _c(vp), nx, ny, nz)
    return np.stack([nx, ny, nz], 1)


def normal_from_points_f64(p):
    p = np.asarray(p, np.float64)
    nv = np.zeros(3, np.float64)
    ev = np.zeros(3, np.float64)
    lib().orc_normal_from_points_f64(_c(p[:, 0], np.float64), _c(p[:, 1], np.float64),
                                     _c(p[:, 2], np.float64), len(p), nv, ev)
    return nv, ev


def eigen_e1(cov, iters=4):
    out = np.zeros(3, np.float32)
    lib().orc_eigen_e1(_c(np.asarray(cov).reshape(-1)), iters, out)
    return out


class RandState(C.Structure):
    _fields_ = [("r", C.c_int32 * 34), ("pos", C.c_int)]


def rand_stream(seed, count):
    st = RandState()
    lib().orc_srand(C.byref(st), seed)
    return np.array([lib().orc_rand(C.byref(st)) for _ in range(count)], np.int64)


def plane_model(p1, p2, p3, collinear=1e-6):
    m = np.zeros(4, np.float32)
    ok = lib().orc_plane_model(_c(p1), _c(p2), _c(p3), collinear, m)
    return (m if ok else None)


def plane_count(x, y, z, model, thr):
    return lib().orc_plane_count(_c(x), _c(y), _c(z), len(x), _c(model), thr)


def ransac_plane(x, y, z, thr, max_iters, collinear=1e-6, probability=0.99):
    model = np.zeros(4, np.float32)
    it = C.c_int(0)
    best = lib().orc_ransac_plane(_c(x), _c(y), _c(z), len(x), thr, max_iters, collinear, probability, model,
                                  C.byref(it))
    return best, model, it.value


def plane_extract(x, y, z, model, thr):
    n = len(x)
    inl = np.zeros((max(n, 1), 3), np.float32)
    outl = np.zeros((max(n, 1), 3), np.float32)
    ni, no = _sz(0), _sz(0)
    lib().orc_plane_extract(_c(x), _c(y), _c(z), n, _c(model), thr, inl.reshape(-1), C.byref(ni),
                            outl.reshape(-1), C.byref(no))
    return inl[:ni.value], outl[:no.value]


def _clusters(fn, x, y, z, tol, min_size, max_size):
    n = len(x)
    idx = np.zeros(max(n, 1), np.int32)
    off = np.zeros(n + 2, np.int64)
    m = fn(_c(x), _c(y), _c(z), n, tol, int(min_size), int(max_size), idx, off)
    return [idx[off[i]:off[i + 1]].copy() for i in range(m)]


def euclidean_cluster(x, y, z, tol, min_size=1, max_size=2**31 - 1):
    """Clusters (sorted index arrays) in the order of their smallest index."""
    return _clusters(lib().orc_euclidean_cluster, x, y, z, tol, min_size, max_size)


def ref_euclidean_cluster(x, y, z, tol, min_size=1, max_size=2**31 - 1):
    return _clusters(ref().ref_euclidean_cluster, x, y, z, tol, min_size, max_size)


def ref_load_pcd(path, cap):
    out = np.zeros((max(cap, 1), 3), np.float32)
    n = ref().ref_load_pcd(str(path).encode(), out.reshape(-1), cap)
    return out[:max(n, 0)] if n >= 0 else None


def ref_save_pcd(path, xyz, binary):
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
    ref().ref_save_pcd(str(path).encode(), xyz.reshape(-1), len(xyz), 1 if binary else 0)
