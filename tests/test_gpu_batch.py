"""Batched (segmented) stages: every cloud's result equals the single-cloud entry point's and the oracle's."""
import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


def soa(p):
    return [np.ascontiguousarray(p[:, i]) for i in range(3)]


@pytest.fixture(scope="module")
def clouds():
    sizes = [1500, 1, 0, 37, 2600, 700, 4000]
    scales = [1.0, 1.0, 1.0, 3.0, 7.5, 0.2, 12.0]
    return [scenes.plane_sphere_cloud(n, seed=100 + i, scale=s) if n else np.zeros((0, 3), np.float32)
            for i, (n, s) in enumerate(zip(sizes, scales))], scales


def test_batch_sor_matches_per_cloud_and_oracle(clouds):
    cl, _ = clouds
    b = R.Batch(cl)
    kept, means = R.batch_sor_exact(b, 20, 1.0)
    for c, p in enumerate(cl):
        if len(p) == 0:
            assert len(kept[c]) == 0
            continue
        oout, omeans = O.sor_sc(p, 20, 1.0)
        assert (b.split(means)[c] == omeans).all(), c
        assert kept[c].shape == oout.shape and (kept[c] == oout).all(), c


def test_batch_radius_matches_oracle(clouds):
    cl, scales = clouds
    b = R.Batch(cl)
    radius = np.array([0.02 * s for s in scales], np.float32)
    idx, d2, cnt = R.batch_radius_search(b, radius, 64)
    for c, p in enumerate(cl):
        if len(p) == 0:
            continue
        x, y, z = soa(p)
        oi, od, oc, _ = O.hash_radius(x, y, z, float(radius[c]), x, y, z, float(radius[c]), 64)
        ci, cd, cc = b.split(idx)[c], b.split(d2)[c], b.split(cnt)[c]
        assert (cc == oc).all(), c
        for i in range(len(p)):
            m = oc[i]
            assert (ci[i, :m] == oi[i, :m]).all() and (cd[i, :m] == od[i, :m]).all(), (c, i)


@pytest.mark.parametrize("per_wave", [1, 64])
def test_batch_radius_runs_cross_clouds_and_dense_windows(clouds, per_wave):
    """Runs of consecutive positions that span cloud boundaries (tiny and empty clouds between), one cloud with
    a dense blob whose windows are too large for one LDS staging."""
    cl, scales = clouds
    blob = np.random.default_rng(8).normal(0.3, 0.004, (900, 3)).astype(np.float32)
    cl = [cl[1], np.concatenate([cl[0], blob]), cl[2], cl[3], cl[5]]
    sc = [scales[1], scales[0], scales[2], scales[3], scales[5]]
    b = R.Batch(cl)
    radius = np.array([0.02 * s for s in sc], np.float32)
    old = R._lib.tune("queries_per_wave")
    try:
        R._lib.tune("queries_per_wave", per_wave)
        idx, d2, cnt = R.batch_radius_search(b, radius, 48)
    finally:
        R._lib.tune("queries_per_wave", old)
    for c, p in enumerate(cl):
        if len(p) == 0:
            continue
        x, y, z = soa(p)
        oi, od, oc, _ = O.hash_radius(x, y, z, float(radius[c]), x, y, z, float(radius[c]), 48)
        ci, cd, cc = b.split(idx)[c], b.split(d2)[c], b.split(cnt)[c]
        assert (cc == oc).all(), c
        for i in range(len(p)):
            m = oc[i]
            assert (ci[i, :m] == oi[i, :m]).all() and (cd[i, :m] == od[i, :m]).all(), (c, i)


def test_batch_many_small_clouds_layout_fallback():
    """More clouds than one layout block takes (1,100): the multi-launch layout path, radius and k-NN grids."""
    rng = np.random.default_rng(31)
    cl = [rng.uniform(0, 0.3, (int(rng.integers(5, 40)), 3)).astype(np.float32) for _ in range(1100)]
    b = R.Batch(cl)
    idx, d2, cnt = R.batch_radius_search(b, 0.05, 16)
    kept, means = R.batch_sor_exact(b, 3, 1.0)
    for c in range(0, 1100, 37):
        p = cl[c]
        x, y, z = soa(p)
        oi, od, oc, _ = O.hash_radius(x, y, z, 0.05, x, y, z, 0.05, 16)
        ci, cd, cc = b.split(idx)[c], b.split(d2)[c], b.split(cnt)[c]
        assert (cc == oc).all(), c
        for i in range(len(p)):
            m = oc[i]
            assert (ci[i, :m] == oi[i, :m]).all() and (cd[i, :m] == od[i, :m]).all(), (c, i)
        oout, omeans = O.sor_sc(p, 3, 1.0)
        assert (b.split(means)[c] == omeans).all(), c
        assert kept[c].shape == oout.shape and (kept[c] == oout).all(), c


def test_batch_hot_cells_and_repeats():
    """The batched grid is a counting sort whose order of points within a cell comes from atomics: cells
    holding 1,500 exact duplicates each, in shuffled input order, give the oracle's radius lists and SOR
    output, and repeated calls give identical results."""
    rng = np.random.default_rng(5)
    base = rng.random((3000, 3)).astype(np.float32)
    hot = np.concatenate([base, np.repeat(base[:3], 1500, axis=0)])
    hot = hot[rng.permutation(len(hot))]
    other = rng.random((2000, 3)).astype(np.float32) * np.float32(4.0)
    cl = [hot, other]
    b = R.Batch(cl)
    radius = np.array([0.05, 0.2], np.float32)
    idx, d2, cnt = R.batch_radius_search(b, radius, 64)
    idx2, d22, cnt2 = R.batch_radius_search(b, radius, 64)
    assert (idx == idx2).all() and (d2 == d22).all() and (cnt == cnt2).all()
    kept, means = R.batch_sor_exact(b, 8, 1.0)
    kept2, means2 = R.batch_sor_exact(b, 8, 1.0)
    assert (means == means2).all() and all((a == c).all() for a, c in zip(kept, kept2))
    for c, p in enumerate(cl):
        x, y, z = soa(p)
        oi, od, oc, _ = O.hash_radius(x, y, z, float(radius[c]), x, y, z, float(radius[c]), 64)
        ci, cd, cc = b.split(idx)[c], b.split(d2)[c], b.split(cnt)[c]
        assert (cc == oc).all(), c
        for i in range(len(p)):
            m = oc[i]
            assert (ci[i, :m] == oi[i, :m]).all() and (cd[i, :m] == od[i, :m]).all(), (c, i)
        oout, omeans = O.sor_sc(p, 8, 1.0)
        assert (b.split(means)[c] == omeans).all(), c
        assert kept[c].shape == oout.shape and (kept[c] == oout).all(), c


def test_batch_radius_lists_again(clouds):
    """The batch workload's radius lists against the oracle (a second grid build over the same clouds)."""
    cl, scales = clouds
    b = R.Batch(cl)
    radius = np.array([0.02 * s for s in scales], np.float32)
    idx, d2, cnt = R.batch_radius_search(b, radius, 64)
    for c, p in enumerate(cl):
        if len(p) == 0:
            continue
        x, y, z = soa(p)
        oi, od, oc, _ = O.hash_radius(x, y, z, float(radius[c]), x, y, z, float(radius[c]), 64)
        ci, cd, cc = b.split(idx)[c], b.split(d2)[c], b.split(cnt)[c]
        assert (cc == oc).all(), c
        for i in range(len(p)):
            m = oc[i]
            assert (ci[i, :m] == oi[i, :m]).all() and (cd[i, :m] == od[i, :m]).all(), (c, i)


def test_batch_normals_equal_single_cloud(clouds):
    cl, _ = clouds
    b = R.Batch(cl)
    nrm = R.batch_normals_knn(b, 24)
    for c, p in enumerate(cl):
        if len(p) < 25:
            continue
        single = np.stack(R.normal_estimation_sc(p, 24, 0.0), 1)
        assert (b.split(nrm)[c] == single).all(), c
        ref = O.normals_knn_f64(*soa(p), 24)
        cosang = np.abs((single.astype(np.float64) * ref).sum(1))
        err = np.degrees(np.arccos(np.clip(cosang, -1, 1)))
        assert np.median(err) <= 0.01


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("rscale", [0.15, 0.3])
def test_batch_normals_from_radius_lists_equal_search(clouds, rscale):
    """k-NN normals whose k + 1 nearest are read from the capped radius lists equal the searched ones bit for bit.
    At 0.15 x scale some queries hold fewer than 25 hits (searched) and others more; at 0.3 most exceed the cap."""
    cl, scales = clouds
    b = R.Batch(cl)
    radius = np.array([rscale * s for s in scales], np.float32)
    idx, d2, cnt = R.batch_radius_search(b, radius, 64)
    served = int((cnt >= 25).sum())
    assert 0 < served < b.n
    ref = R.batch_normals_knn(b, 24)
    assert same_bits(R.batch_normals_knn_lists(b, 24, idx, d2, cnt), ref)
    assert same_bits(R.batch_normals_knn_lists(b, 24, idx, None, cnt), ref)


def test_batch_normals_lists_with_foreign_indices_fall_back(clouds):
    """Rows holding an index outside their cloud, and rows narrower than k + 1, are searched instead."""
    cl, scales = clouds
    b = R.Batch(cl)
    radius = np.array([0.3 * s for s in scales], np.float32)
    idx, d2, cnt = R.batch_radius_search(b, radius, 64)
    ref = R.batch_normals_knn(b, 24)
    rows = np.flatnonzero(cnt >= 25)[:200]
    bad = idx.copy()
    bad[rows[::2], 3] = -7
    bad[rows[1::2], 24] = 1 << 30
    assert same_bits(R.batch_normals_knn_lists(b, 24, bad, d2, cnt), ref)
    assert same_bits(R.batch_normals_knn_lists(b, 24, idx[:, :16], None, np.minimum(cnt, 16)), ref)


def test_batch_voxel_matches_oracle(clouds):
    cl, scales = clouds
    b = R.Batch(cl)
    leaf = np.array([0.01 * s for s in scales], np.float32)
    vox, status = R.batch_voxel(b, leaf)
    assert (status == 0).all()
    for c, p in enumerate(cl):
        ref = O.voxel_sc(p, float(leaf[c])) if len(p) else np.zeros((0, 3), np.float32)
        assert vox[c].shape == ref.shape and (vox[c] == ref).all(), c


def test_batch_voxel_wide_and_negative_keys():
    rng = np.random.default_rng(7)
    a = (rng.random((3000, 3)) * 400 - 200).astype(np.float32)   # wide range: keys above 2^32 at leaf 0.01
    bcl = (rng.random((500, 3)) - 0.5).astype(np.float32)
    b = R.Batch([a, bcl])
    vox, status = R.batch_voxel(b, [0.01, 0.05])
    assert (status == 0).all()
    for c, (p, lf) in enumerate([(a, 0.01), (bcl, 0.05)]):
        ref = O.voxel_sc(p, lf)
        assert (vox[c] == ref).all()


@pytest.mark.parametrize("per_wave", [0, 3, 64])
def test_batch_normals_radius_equal_single_cloud(clouds, per_wave):
    """Batched radius normals with 8 (default), 3 or 64 consecutive queries per wave (batches that share staged
    windows, cloud boundaries inside them) equal the single-cloud call at default settings."""
    cl, scales = clouds
    b = R.Batch(cl)
    radius = np.array([0.03 * s for s in scales], np.float32)
    old = R._lib.tune("normals_per_wave")
    try:
        R._lib.tune("normals_per_wave", per_wave)
        nrm = R.batch_normals_radius(b, radius, 32)
    finally:
        R._lib.tune("normals_per_wave", old)
    for c, p in enumerate(cl):
        if len(p) == 0:
            continue
        h = R.SpatialHash()
        h.setInputCloud(p, float(radius[c]))
        h.build()
        single = np.stack(R.normal_estimation_rvv(p, h, 10, float(radius[c]), max_nn=32), 1)
        assert (b.split(nrm)[c] == single).all(), c


def test_radius_normals_independent_of_path_and_order(clouds):
    """Radius normals read the neighbour set, not its order: a query with at most max_nn hits takes the unsorted
    staged path (its neighbours in the batched grid's within-cell order, which comes from atomics and varies
    between calls), one with more takes the sorted top-K; both sum exactly (terms rounded to a common quantum), so
    repeated calls agree bit for bit, and a query whose hit count is within both caps gets the same normal from a
    call capped at 24 and one capped at 64 (one path or the other, the same neighbour set)."""
    cl, scales = clouds
    b = R.Batch(cl)
    radius = np.array([0.06 * s for s in scales], np.float32)
    first = R.batch_normals_radius(b, radius, 64)
    for _ in range(3):
        assert (R.batch_normals_radius(b, radius, 64) == first).all()
    capped = R.batch_normals_radius(b, radius, 24)
    n_small = n_large = 0
    for c, p in enumerate(cl):
        if len(p) < 2:
            continue
        x, y, z = soa(p)
        _, _, cnt, _ = O.hash_radius(x, y, z, float(radius[c]), x, y, z, float(radius[c]), 64)
        small = cnt <= 24
        n_small += int(small.sum())
        n_large += int((cnt > 24).sum())
        assert (b.split(capped)[c][small] == b.split(first)[c][small]).all(), c
    assert n_small > 100 and n_large > 100  # both paths were taken


def test_radius_normals_unsorted_path_equals_sorted_path():
    """The same neighbour sets through both paths of the radius normals: with the grid's cell equal to the radius,
    a 3 x 3 x 3 window of this uniform volume holds about 320 points, too many to stage, so every query walks it
    and keeps the sorted top-K; with half that cell the window (about 190 points) is staged and the queries with
    at most 64 hits (about 50 on average) take the unsorted path.  The covariances sum exactly either way, so
    the normals are equal bit for bit."""
    rng = np.random.default_rng(7)
    p = rng.random((20000, 3)).astype(np.float32)
    r = float(np.cbrt(12.0 / 20000.0))
    b = R.Batch([p])
    x, y, z = soa(p)
    _, _, cnt, _ = O.hash_radius(x, y, z, r, x, y, z, r, 64)
    assert np.mean(cnt < 64) > 0.8
    walked = R.batch_normals_radius(b, r, 64, cell=r)
    staged = R.batch_normals_radius(b, r, 64, cell=0.5 * r)
    assert (walked == staged).all()


def test_batch_ransac_matches_single_and_oracle(clouds):
    cl, scales = clouds
    b = R.Batch(cl)
    models, inl, its = R.batch_ransac_plane(b, 0.005 * 5, 300)
    for c, p in enumerate(cl):
        best, model = R.ransac_plane_sc(p, 0.005 * 5, 300)
        assert inl[c] == best and (models[c] == model).all(), c
        if len(p) >= 3:
            ob, om, oit = O.ransac_plane(*soa(p), 0.005 * 5, 300)
            assert ob == inl[c] and (om == models[c]).all() and oit == its[c], c


@pytest.mark.parametrize("nb,count", [(5000, 0), (5000, 1), (5000, 777), (5000, 5000), (5000, 6000),
                                      (300_000, 290_000)])
def test_aos_to_soa_bounded_reads_the_device_count(nb, count):
    """Converts min(*n_dev, n_bound) points (grid-stride beyond 262,144) and leaves the rest untouched."""
    from rvpoint_amd._lib import call
    rng = np.random.default_rng(count)
    aos = rng.random((nb, 3)).astype(np.float32)
    D = R.DeviceArray
    a = D.from_numpy(aos)
    cnt = D.from_numpy(np.array([count], np.uint32))
    outs = [D.from_numpy(np.full(nb, -1.0, np.float32)) for _ in range(3)]
    call("rvp_aos_to_soa_bounded", a.ptr, nb, cnt.ptr, *(o.ptr for o in outs), None)
    m = min(count, nb)
    for j, o in enumerate(outs):
        got = o.numpy()
        assert (got[:m] == aos[:m, j]).all() and (got[m:] == -1.0).all(), j


def test_batch_chained_stages_without_host_sync():
    """voxel -> SOR -> radius normals -> RANSAC on device offsets equals the per-cloud chain."""
    import ctypes as C

    from rvpoint_amd._lib import call
    cl = [scenes.plane_sphere_cloud(n, seed=200 + i, scale=s) for i, (n, s) in
          enumerate([(3000, 1.0), (800, 2.0), (5000, 4.0)])]
    b = R.Batch(cl)
    n, ncl = b.n, b.ncl
    D = R.DeviceArray
    leaf = D.from_numpy(np.array([0.01, 0.02, 0.04], np.float32))
    vox, voff, st = D((n, 3)), D(ncl + 1, np.uint32), D(ncl, np.uint32)
    call("rvp_batch_voxel", *b.dev.ptrs(), b.offsets.ptr, ncl, n, leaf.ptr, vox.ptr, voff.ptr, st.ptr, None)
    vx, vy, vz = D(n), D(n), D(n)
    call("rvp_aos_to_soa", vox.ptr, n, vx.ptr, vy.ptr, vz.ptr, None)
    sor, soff = D((n, 3)), D(ncl + 1, np.uint32)
    call("rvp_batch_sor_exact", vx.ptr, vy.ptr, vz.ptr, voff.ptr, ncl, n, 10, 1.0, sor.ptr, soff.ptr, None, None,
         None)
    sx, sy, sz = D(n), D(n), D(n)
    call("rvp_aos_to_soa", sor.ptr, n, sx.ptr, sy.ptr, sz.ptr, None)
    rad = D.from_numpy(np.array([0.05, 0.1, 0.2], np.float32))
    nx, ny, nz = D(n), D(n), D(n)
    call("rvp_batch_normals_radius", sx.ptr, sy.ptr, sz.ptr, soff.ptr, ncl, n, rad.ptr, None, 64, 0.0, 0.0, 0.0,
         nx.ptr, ny.ptr, nz.ptr, None)
    m = np.zeros((ncl, 4), np.float32)
    inl = np.zeros(ncl, np.int32)
    its = np.zeros(ncl, np.int32)
    call("rvp_batch_ransac_plane", sx.ptr, sy.ptr, sz.ptr, soff.ptr, ncl, n, 0.02, 200, 1e-6, 0.99,
         m.ctypes.data_as(C.POINTER(C.c_float)), inl.ctypes.data_as(C.POINTER(C.c_int)),
         its.ctypes.data_as(C.POINTER(C.c_int)), None)
    so = soff.numpy()
    S = sor.numpy()
    N = np.stack([nx.numpy(), ny.numpy(), nz.numpy()], 1)
    for c, (p, lf, r) in enumerate(zip(cl, [0.01, 0.02, 0.04], [0.05, 0.1, 0.2])):
        v = O.voxel_sc(p, lf)
        kept, _ = O.sor_sc(v, 10, 1.0)
        assert (S[so[c]:so[c + 1]] == kept).all(), c
        h = R.SpatialHash()
        h.setInputCloud(kept, r)
        h.build()
        single = np.stack(R.normal_estimation_rvv(kept, h, 10, r, max_nn=64), 1)
        assert (N[so[c]:so[c + 1]] == single).all(), c
        best, model = R.ransac_plane_sc(kept, 0.02, 200)
        assert inl[c] == best and (m[c] == model).all(), c
