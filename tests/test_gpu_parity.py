"""GPU parity: every HIP entry point against the CPU oracle on the same seeded inputs.

Bit-exact: neighbour lists (indices, d2, order, counts), voxel centroids, SOR keep-sets and means, RANSAC
count/model/iterations, plane extraction.  Normals: angular error against the FP64 oracle on the same
neighbourhoods, median <= 0.01 deg (sign-agnostic), p99 <= 1 deg."""
import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


def soa(p):
    return [np.ascontiguousarray(p[:, i]) for i in range(3)]


def gauss(n, seed, scale=1.0):
    return np.random.default_rng(seed).normal(0, scale, (n, 3)).astype(np.float32)


def c0_small(n=3000, seed=7):
    return scenes.plane_sphere_cloud(n, seed=seed)


# ------------------------------------------------------------------------------------------------ H1/H2
@pytest.mark.parametrize("cell", [0.02, 0.1, 0.37])
def test_hash_bbox_and_cells(cell):
    p = c0_small()
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), cell)
    h.build()
    assert (h.bbox() == O.bbox(x, y, z, cell)).all()
    ixyz, keys = h.cell_coords()
    oi, ok = O.cell_coords(x, y, z, cell)
    assert (ixyz == oi).all() and (keys == ok).all()


# ------------------------------------------------------------------------------------------------ H3
def _check_lists(got, idx, d2, cnt):
    for i, (gi, gd) in enumerate(got):
        m = cnt[i]
        assert len(gi) == m, i
        assert (gi == idx[i, :m]).all(), i
        assert (gd == d2[i, :m]).all(), i


@pytest.mark.parametrize("cell,r,max_nn", [(0.02, 0.02, 64), (0.02, 0.05, 16), (0.05, 0.02, 1), (0.01, 0.03, 64),
                                           (0.02, 0.02, 100)])
def test_radius_capped_external_queries(cell, r, max_nn):
    p = c0_small(4000, 11)
    x, y, z = soa(p)
    rng = np.random.default_rng(3)
    q = np.concatenate([p[:300], rng.uniform(-0.1, 1.1, (100, 3)).astype(np.float32)])
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), cell)
    h.build()
    got = h.radius_search_batch(q, r, max_nn)
    idx, d2, cnt, _ = O.hash_radius(x, y, z, cell, *soa(q), r, max_nn)
    _check_lists(got, idx, d2, cnt)


def dense_blob_cloud():
    """c0-like surfaces plus a blob of 1200 points within ~0.02 of one spot: its windows hold far more points
    than one LDS staging takes, the rest stage; consecutive points share windows."""
    blob = np.random.default_rng(5).normal(0.5, 0.006, (1200, 3)).astype(np.float32)
    return np.concatenate([c0_small(4000, 21), blob])


@pytest.mark.parametrize("per_wave", [0, 1, 7, 64])
def test_radius_capped_runs_of_queries(per_wave):
    """Capped lists with each wave serving runs of consecutive queries (knob forced; 0 = from the count), over
    windows that stage in LDS and windows too large to stage; the cloud's own points and external queries."""
    p = dense_blob_cloud()
    x, y, z = soa(p)
    rng = np.random.default_rng(9)
    q = np.concatenate([p[::3], rng.uniform(-0.1, 1.1, (200, 3)).astype(np.float32)])
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.02)
    h.build()
    old = R._lib.tune("queries_per_wave")
    try:
        R._lib.tune("queries_per_wave", per_wave)
        # r = 0.2 over 0.02 cells: windows of more than 64 (y, z) rows, walked without staging
        for r, mnn in ((0.02, 64), (0.03, 17), (0.2, 8)):
            got = h.radius_search_batch(None, r, mnn)
            idx, d2, cnt, _ = O.hash_radius(x, y, z, 0.02, x, y, z, r, mnn)
            _check_lists(got, idx, d2, cnt)
            assert cnt.max() == mnn  # the blob's lists are truncated
            got = h.radius_search_batch(q, r, mnn)
            idx, d2, cnt, _ = O.hash_radius(x, y, z, 0.02, *soa(q), r, mnn)
            _check_lists(got, idx, d2, cnt)
    finally:
        R._lib.tune("queries_per_wave", old)


@pytest.mark.parametrize("cell,r", [(0.02, 0.02), (0.05, 0.08)])
def test_radius_unlimited_self(cell, r):
    p = c0_small(3000, 12)
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), cell)
    h.build()
    got = h.radius_search_batch(None, r, 0)
    idx, d2, cnt, _ = O.hash_radius(x, y, z, cell, x, y, z, r, 0)
    _check_lists(got, idx, d2, cnt)


def test_radius_long_lists_lds_and_global_sort():
    """rows of 65..4096 (LDS sort) and > 4096 (global-memory sort)."""
    p = gauss(9000, 13, 0.3)
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.2)
    h.build()
    q = np.array([[0, 0, 0], [0.2, 0.1, 0.0], [1.0, 1.0, 1.0]], np.float32)
    for r in (0.15, 0.5, 2.0):
        got = h.radius_search_batch(q, r, 0)
        idx, d2, cnt, _ = O.radius_all(x, y, z, *soa(q), r, 0)
        _check_lists(got, idx, d2, cnt)
    assert max(len(g[0]) for g in got) > 4096


def test_radius_duplicates_tie_break_by_index():
    base = gauss(50, 14, 0.01)
    p = np.repeat(base, 30, axis=0)          # exact duplicates: equal d2, order by index
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.01)
    h.build()
    for mnn in (0, 20, 64):
        got = h.radius_search_batch(p[:40], 0.01, mnn)
        idx, d2, cnt, _ = O.hash_radius(x, y, z, 0.01, *soa(p[:40]), 0.01, mnn)
        _check_lists(got, idx, d2, cnt)


def test_radius_empty_and_tiny():
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(np.zeros(0), np.zeros(0), np.zeros(0)), 0.1)
    h.build()
    assert [len(g[0]) for g in h.radius_search_batch([[0, 0, 0]], 1.0, 5)] == [0]
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA([1.0], [2.0], [3.0]), 0.1)
    h.build()
    i, d = h.radiusSearch([1.0, 2.0, 3.0], 0.5)
    assert list(i) == [0] and list(d) == [0.0]


def test_radius_reference_unit_test_shape():
    """tests/unit/test_radius.cpp input shape: 1000 N(0,2) points, q=0, r=2, max_nn=100: brute force sc/rvv."""
    p = gauss(1000, 42, 2.0)
    for mnn in (100, 5, 0):
        gi, gd = R.radius_search_sc(p, (0, 0, 0), 2.0, mnn)
        oi, od = O.radius_bruteforce(*soa(p), (0.0, 0.0, 0.0), 2.0, mnn)
        assert (gi == oi).all() and (gd == od).all()
        gi2, gd2 = R.radius_search_rvv(R.PointCloudSoA(*soa(p)), (0, 0, 0), 2.0, mnn)
        assert (gi2 == oi).all() and (gd2 == od).all()


# ------------------------------------------------------------------------------------------------ kNN
@pytest.mark.parametrize("k", [1, 11, 33, 64])
def test_knn(k):
    p = c0_small(2500, 15)
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.03)
    h.build()
    q = np.concatenate([p[:200], np.random.default_rng(4).uniform(-0.5, 1.5, (50, 3)).astype(np.float32)])
    gi, gd = h.knn(q, k)
    oi, od = O.knn(x, y, z, *soa(q), k)
    assert (gi == oi).all() and (gd == od).all()


def test_knn_fewer_points_than_k():
    p = gauss(20, 16)
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.5)
    h.build()
    gi, gd = h.knn(p, 32)
    oi, od = O.knn(x, y, z, x, y, z, 32)
    assert (gi == oi).all() and (gd[:, :20] == od[:, :20]).all() and np.isinf(gd[:, 20:]).all()


# ------------------------------------------------------------------------------------------------ SOR
@pytest.mark.parametrize("k,alpha", [(10, 1.0), (32, 1.0), (50, 0.5), (20, 2.0)])
def test_sor_exact(k, alpha):
    p = c0_small(3000, 17)
    out, means = R.sor_sc(p, k, alpha, with_means=True)
    oout, omeans = O.sor_sc(p, k, alpha)
    assert (means == omeans).all()
    assert out.shape == oout.shape and (out == oout).all()
    out2 = R.sor_rvv(R.PointCloudSoA(*soa(p)), k, alpha)
    assert (out2 == oout).all()


@pytest.mark.parametrize("case", ["offset", "duplicates", "tiny_steps"])
def test_sor_exact_previous_query_bound(case):
    """Exact k-NN means where the bound from the previous query is tight or void: coordinates far from the
    origin (large roundings in the distances), runs of duplicate points (k-th distance 0: no bound), and a
    dense strand whose consecutive queries are nearly coincident; single-cloud and batched paths."""
    rng = np.random.default_rng(["offset", "duplicates", "tiny_steps"].index(case) + 70)
    if case == "offset":
        p = (c0_small(3000, 19).astype(np.float64) * 0.01 + [1.0e5, -3.0e4, 2.5e3]).astype(np.float32)
    elif case == "duplicates":
        base = c0_small(1500, 20)
        p = np.concatenate([base, np.repeat(base[:40], 30, axis=0), base[::7]]).astype(np.float32)
    else:
        t = np.sort(rng.random(2500))
        p = np.stack([t, np.sin(7 * t) * 0.01, 1e-6 * rng.random(2500)], 1).astype(np.float32)
        p = np.concatenate([p, c0_small(800, 21)]).astype(np.float32)
    for k in (5, 30):
        out, means = R.sor_sc(p, k, 1.0, with_means=True)
        oout, omeans = O.sor_sc(p, k, 1.0)
        assert (means == omeans).all(), k
        assert out.shape == oout.shape and (out == oout).all(), k
        b = R.Batch([p, p[: len(p) // 3]])
        kept, bmeans = R.batch_sor_exact(b, k, 1.0)
        assert (b.split(bmeans)[0] == omeans).all(), k
        assert (kept[0] == oout).all(), k


def test_sor_reference_unit_test_shape():
    """tests/unit/test_sor.cpp input shape: 900 N(0,1) + 100 U(-20,20), k=10, alpha=1."""
    rng = np.random.default_rng(42)
    p = np.concatenate([rng.normal(0, 1, (900, 3)), rng.uniform(-20, 20, (100, 3))]).astype(np.float32)
    out = R.sor_sc(p, 10, 1.0)
    oout, _ = O.sor_sc(p, 10, 1.0)
    assert (out == oout).all()


@pytest.mark.parametrize("cell,r,k", [(0.02, 0.02, 10), (0.05, 0.05, 32), (0.5, 0.5, 20)])
def test_sor_radius(cell, r, k):
    p = c0_small(4000, 18)
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), cell)
    h.build()
    out, means = R.sor_spatial_hash(R.PointCloudSoA(x, y, z), h, k, 1.0, r, with_means=True)
    oout, omeans = O.sor_radius(x, y, z, cell, k, 1.0, r)
    assert (means == omeans).all()
    assert (out == oout).all()


# ------------------------------------------------------------------------------------------------ V1
@pytest.mark.parametrize("leaf", [0.01, 0.05, 0.3])
def test_voxel(leaf):
    p = c0_small(20000, 19)
    v = R.voxel_grid_downsamp_sc(p, leaf)
    ov = O.voxel_sc(p, leaf)
    assert v.shape == ov.shape and (v == ov).all()


def test_voxel_negative_and_wide_keys():
    rng = np.random.default_rng(20)
    p = np.concatenate([rng.uniform(-500, 500, (3000, 3)), rng.normal(0, 1, (3000, 3))]).astype(np.float32)
    for leaf in (0.001, 0.5):             # 0.001 needs 64-bit keys
        v = R.voxel_grid_downsamp_rvv_v2(R.PointCloudSoA(*soa(p)), leaf)
        assert (v == O.voxel_sc(p, leaf)).all()


def test_voxel_kat_kitti():
    import os
    xyz = np.load(os.path.join(os.path.dirname(__file__), "golden", "kitti_0000000000.npz"))["xyz"]
    v = R.voxel_grid_downsamp_sc(xyz, 0.2)
    assert len(v) == 18542
    assert (v == O.voxel_sc(xyz, 0.2)).all()


# ------------------------------------------------------------------------------------------------ R1 / X1
@pytest.mark.parametrize("seed", [42, 1, 2])
def test_ransac(seed):
    rng = np.random.default_rng(seed)
    p = np.concatenate([np.c_[rng.uniform(-10, 10, (500, 2)), rng.normal(10, 0.01, 500)],
                        rng.uniform(-20, 20, (500, 3))]).astype(np.float32)
    best, model = R.ransac_plane_sc(p, 0.1, 1000)
    ob, om, _ = O.ransac_plane(*soa(p), 0.1, 1000)
    assert best == ob and (model == om).all()
    assert best >= 450 and abs(model[2]) > 0.9 and abs(model[3]) >= 9.0


def test_ransac_large_max_iters_and_low_inlier_ratio():
    p = c0_small(20000, 21)
    for thr, iters in [(0.005, 500), (0.001, 3000)]:
        best, model = R.ransac_plane_rvv(R.PointCloudSoA(*soa(p)), thr, iters)
        ob, om, _ = O.ransac_plane(*soa(p), thr, iters)
        assert best == ob and (model == om).all()


def test_plane_extract():
    p = c0_small(5000, 22)
    m = np.array([0.0, 0.0, 1.0, -0.2], np.float32)
    a, b = R.extract_plane_inliers_outliers_rvv(R.PointCloudSoA(*soa(p)), m, 0.005)
    oa, ob = O.plane_extract(*soa(p), m, 0.005)
    assert (a == oa).all() and (b == ob).all()


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 100000, 1000003])
def test_plane_split_one_pass_sizes(n):
    """The one-pass split at tile edges and past the look-back window (489 tiles at 10^6 points): both outputs
    against the oracle, NaN coordinates on the outlier side, and each output alone."""
    import ctypes as C

    from rvpoint_amd import _lib
    rng = np.random.default_rng(n)
    p = np.c_[rng.uniform(-5, 5, (n, 2)), rng.normal(0, 0.05, n)].astype(np.float32)
    p[::97, 2] = np.nan
    m = np.array([0.0, 0.0, 1.0, 0.0], np.float32)
    print(f"plane split n={n}", flush=True)
    a, b = R.extract_plane_inliers_outliers_rvv(R.PointCloudSoA(*soa(p)), m, 0.05)
    oa, ob = O.plane_extract(*soa(p), m, 0.05)
    assert len(a) + len(b) == n
    assert np.array_equal(a.view(np.uint32), oa.view(np.uint32)) and np.array_equal(b.view(np.uint32),
                                                                                     ob.view(np.uint32))
    dc = R.DeviceCloud.from_soa(*soa(p))
    out = R.DeviceArray((n, 3), np.float32)
    mm = (C.c_float * 4)(*[float(v) for v in m])
    ni, no = C.c_size_t(), C.c_size_t()
    _lib.call("rvp_plane_extract", *dc.ptrs(), n, mm, 0.05, None, C.byref(ni), out.ptr, C.byref(no), None)
    assert ni.value == len(oa) and np.array_equal(out.numpy()[:no.value].view(np.uint32), ob.view(np.uint32))
    _lib.call("rvp_plane_extract", *dc.ptrs(), n, mm, 0.05, out.ptr, C.byref(ni), None, C.byref(no), None)
    assert no.value == len(ob) and np.array_equal(out.numpy()[:ni.value].view(np.uint32), oa.view(np.uint32))


@pytest.mark.parametrize("seed", [0, 1])
def test_ransac_plane_extract_fused(seed):
    """RANSAC and the split by its best model in one call equal the two reference steps in sequence."""
    p = c0_small(30000, 40 + seed)
    best, model, iters, a, b = R.ransac_plane_extract(R.PointCloudSoA(*soa(p)), 0.005, 500)
    ob, om, oit = O.ransac_plane(*soa(p), 0.005, 500)
    assert best == ob and (model == om).all() and iters == oit
    oa, obb = O.plane_extract(*soa(p), om, 0.005)
    assert (a == oa).all() and (b == obb).all() and len(a) + len(b) == len(p)


# ------------------------------------------------------------------------------------------------ normals
def _angles_deg(n, ref):
    n = np.asarray(n, np.float64)
    ref = np.asarray(ref, np.float64)
    c = np.abs((n * ref).sum(1)) / np.maximum(np.linalg.norm(n, axis=1) * np.linalg.norm(ref, axis=1), 1e-300)
    return np.degrees(np.arccos(np.clip(c, -1, 1)))


@pytest.mark.parametrize("k", [10, 24, 30])
def test_normals_knn_vs_fp64(k):
    p = c0_small(3000, 23)
    nx, ny, nz = R.normal_estimation_sc(p, k, 0.0)
    ref = O.normals_knn_f64(*soa(p), k)
    err = _angles_deg(np.c_[nx, ny, nz], ref)
    assert np.median(err) <= 0.01 and np.percentile(err, 99) <= 1.0, (np.median(err), np.percentile(err, 99))
    # flip rule toward the viewpoint (origin), evaluated in fp64
    dots = (-p.astype(np.float64) * np.c_[nx, ny, nz]).sum(1)
    assert (dots >= -1e-5).all()


@pytest.mark.parametrize("r,max_nn", [(0.03, 64), (0.03, 0), (0.05, 16), (0.02, 100)])
def test_normals_radius_vs_fp64(r, max_nn):
    p = c0_small(3000, 24)
    x, y, z = soa(p)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), r)
    h.build()
    nx, ny, nz = R.normal_estimation_rvv(R.PointCloudSoA(x, y, z), h, 10, r, max_nn=max_nn)
    ref = O.normals_radius_f64(x, y, z, r, r, max_nn)
    zero_ref = np.linalg.norm(ref, axis=1) == 0
    got = np.c_[nx, ny, nz]
    assert ((np.linalg.norm(got, axis=1) == 0) == zero_ref).all()
    err = _angles_deg(got[~zero_ref], ref[~zero_ref])
    assert np.median(err) <= 0.01 and np.percentile(err, 99) <= 1.0, (np.median(err), np.percentile(err, 99))


def test_normals_reference_unit_test_shape():
    """tests/unit/test_normal.cpp: plane z ~ N(0, 0.01) over [-10,10]^2, k=10: >= 95% with |nz| > 0.9."""
    rng = np.random.default_rng(42)
    p = np.c_[rng.uniform(-10, 10, (1000, 2)), rng.normal(0, 0.01, 1000)].astype(np.float32)
    nx, ny, nz = R.normal_estimation_sc(p, 10, 2.0)
    assert (np.abs(nz) > 0.9).mean() >= 0.95


# ------------------------------------------------------------------------------------------------ exact sums
def _gpu_seq_sum(v, squared=False, mean=0.0):
    import ctypes as C
    from rvpoint_amd._lib import call
    dv = R.DeviceArray.from_numpy(np.ascontiguousarray(v, np.float32))
    dm = R.DeviceArray.from_numpy(np.array([mean], np.float32))
    out = R.DeviceArray(1, np.float32)
    call("rvp_seq_sum", dv.ptr, len(v), 1 if squared else 0, dm.ptr, out.ptr, None)
    return out.numpy()[0]


SUM_KINDS = ["uniform", "ties", "wide", "zeros_first", "tiny", "large", "nan", "growing", "sparse", "subnormal",
             "inf", "signed"]


@pytest.mark.parametrize("kind", SUM_KINDS)
def test_exact_sequential_sum(kind):
    rng = np.random.default_rng(SUM_KINDS.index(kind))
    if kind == "uniform":
        v = rng.random(200_003)
    elif kind == "ties":      # halfway cases at every binade: exercises round-to-even on the running sum
        v = rng.integers(0, 1000, 300_000) * 2.0 ** -14
    elif kind == "wide":
        v = np.exp(rng.normal(0, 4, 100_000))
    elif kind == "zeros_first":
        v = np.concatenate([np.zeros(5000), rng.random(50_000) * 1e-30, rng.random(50_000)])
    elif kind == "tiny":
        v = rng.random(37)
    elif kind == "large":
        v = rng.random(3_000_000) * 0.1
    elif kind == "growing":   # the sum crosses a binade every few hundred terms, inside chunks and lanes
        v = 1.02 ** np.arange(1500) * (1 + rng.random(1500))
    elif kind == "sparse":    # long runs of zeros, also at the start of the sum
        v = np.where(rng.random(120_000) < 0.01, rng.random(120_000), 0.0)
        v[:700] = 0.0
    elif kind == "subnormal":  # a subnormal running sum for a while, then normal
        v = np.concatenate([np.full(3000, 2.0 ** -140), rng.random(20_000) * 2.0 ** -100, rng.random(5000)])
    elif kind == "inf":
        v = rng.random(20_000)
        v[12_345] = np.inf
    elif kind == "signed":    # outside the SOR contract (terms >= 0); the sum is still the fp32 adder's
        v = rng.normal(0, 1, 30_000)
    else:
        v = rng.random(10_000)
        v[7777] = np.nan
    v = v.astype(np.float32)
    got = _gpu_seq_sum(v)
    exp = O.seq_sum(v)
    assert (np.isnan(got) and np.isnan(exp)) or got == exp
    if kind not in ("nan", "inf"):
        mean = np.float32(exp / np.float32(len(v)))
        assert _gpu_seq_sum(v, True, mean) == O.seq_sum(v, True, mean)



@pytest.mark.parametrize("n", [1, 5, 2048, 2049, 130_000, 1_000_007, 4_194_304])
def test_single_pass_scan(n):
    """The single-pass (look-back) exclusive scan against numpy, out of place and in place, with the total."""
    import ctypes as C

    from rvpoint_amd import _lib
    rng = np.random.default_rng(n)
    v = rng.integers(0, 300, n, dtype=np.uint32)
    d_in = R.DeviceArray.from_numpy(v)
    d_out = R.DeviceArray(n, np.uint32)
    tot = R.DeviceArray(1, np.uint32)
    _lib.call("rvp_exclusive_scan_u32", d_in.ptr, d_out.ptr, n, tot.ptr, None)
    ref = np.concatenate([[0], np.cumsum(v, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert (d_out.numpy() == ref).all() and int(tot.numpy()[0]) == int(v.sum(dtype=np.uint64)) & 0xffffffff
    _lib.call("rvp_exclusive_scan_u32", d_in.ptr, d_in.ptr, n, None, None)   # in place
    assert (d_in.numpy() == ref).all()
    errs = C.c_ulonglong()
    _lib.call("rvp_scan_wait_errors", C.byref(errs))
    assert errs.value == 0


def _sort_case(dtype, n, bits, n_dev=None, bits_dev=None, seed=0, ordered=False):
    import ctypes as C

    from rvpoint_amd import _lib
    rng = np.random.default_rng(seed)
    top = (1 << 64) - 1 if dtype == np.uint64 else (1 << 32) - 1
    keys = rng.integers(0, top, n, dtype=dtype, endpoint=True)
    if ordered:  # long runs of equal high digits, as cell keys of scan-ordered points have
        keys = np.sort(keys >> dtype(24)).astype(dtype) | (keys & dtype(0xff))
    vals = np.arange(n, dtype=np.uint32)
    dk, dv = R.DeviceArray.from_numpy(keys), R.DeviceArray.from_numpy(vals)
    tk, tv = R.DeviceArray(n, dtype), R.DeviceArray(n, np.uint32)
    nd = R.DeviceArray.from_numpy(np.array([n_dev], np.uint32)) if n_dev is not None else None
    bd = R.DeviceArray.from_numpy(np.array([bits_dev], np.int32)) if bits_dev is not None else None
    in_tmp = C.c_int(-1)
    name = "rvp_sort_pairs_u64" if dtype == np.uint64 else "rvp_sort_pairs_u32"
    print(f"sort case {name} n={n} bits={bits} n_dev={n_dev} bits_dev={bits_dev} ordered={ordered}", flush=True)
    _lib.call(name, dk.ptr, dv.ptr, tk.ptr, tv.ptr, n, bits, nd.ptr if nd else None, bd.ptr if bd else None,
              C.byref(in_tmp), None)
    ok, ov = (tk, tv) if in_tmp.value else (dk, dv)
    m = n if n_dev is None else min(n_dev, n)
    eff = bits if bits_dev is None else min(bits, bits_dev)
    sort_bits = min(8 * ((eff + 7) // 8), 8 * np.dtype(dtype).itemsize)
    mask = dtype(top) if sort_bits >= 8 * np.dtype(dtype).itemsize else dtype((1 << sort_bits) - 1)
    order = np.argsort(keys[:m] & mask, kind="stable")
    assert np.array_equal(ov.numpy()[:m], vals[:m][order])
    assert np.array_equal(ok.numpy()[:m], keys[:m][order])


@pytest.mark.parametrize("n", [2, 1000, 2047, 2048, 2049, 100003, 1 << 20])
@pytest.mark.parametrize("bits", [8, 17, 32])
def test_radix_sort_u32(n, bits):
    """The one-sweep radix sort: keys and values against numpy's stable argsort on the sorted digits."""
    _sort_case(np.uint32, n, bits, seed=n + bits)


@pytest.mark.parametrize("n_dev,bits_dev", [(None, None), (None, 0), (None, 20), (None, 40), (None, 64),
                                            (333333, None), (333333, 33), (0, None), (1, 64)])
def test_radix_sort_u64_device_sized(n_dev, bits_dev):
    _sort_case(np.uint64, 1000003, 64, n_dev, bits_dev, seed=7)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_radix_sort_ordered_input_and_wait_errors(dtype):
    import ctypes as C

    from rvpoint_amd import _lib
    _sort_case(dtype, 1 << 21, 8 * np.dtype(dtype).itemsize, seed=3, ordered=True)
    errs = C.c_ulonglong()
    _lib.call("rvp_scan_wait_errors", C.byref(errs))
    assert errs.value == 0
