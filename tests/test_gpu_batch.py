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
