"""Euclidean clustering on the GPU against the oracle (the reference's breadth-first search restated):
identical clusters, in the same order, with size filters, duplicates and degenerate inputs."""
import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


def soa(p):
    return [np.ascontiguousarray(p[:, i]) for i in range(3)]


def same(a, b):
    return len(a) == len(b) and all(np.array_equal(u, v) for u, v in zip(a, b))


@pytest.mark.parametrize("tol", [0.01, 0.02, 0.05])
def test_clusters_match_oracle(tol):
    p = scenes.plane_sphere_cloud(6000, seed=91)
    assert same(R.euclidean_cluster(p, tol), O.euclidean_cluster(*soa(p), tol))


def test_size_filters_and_order():
    rng = np.random.default_rng(5)
    blobs = [rng.normal(c, s, (m, 3)) for c, s, m in [((0, 0, 0), 0.02, 400), ((3, 0, 0), 0.02, 60),
                                                     ((0, 3, 0), 0.5, 300), ((5, 5, 5), 0.01, 5)]]
    p = np.concatenate(blobs).astype(np.float32)
    p = p[rng.permutation(len(p))]
    for mn, mx in [(1, 1 << 30), (10, 1 << 30), (50, 350), (1000, 2000)]:
        assert same(R.euclidean_cluster(p, 0.08, mn, mx), O.euclidean_cluster(*soa(p), 0.08, mn, mx)), (mn, mx)


def test_duplicates_zero_tolerance_and_tiny_clouds():
    p = np.array([[0, 0, 0], [1, 1, 1], [0, 0, 0], [1, 1, 1.0000001], [0, 0, 0]], np.float32)
    assert same(R.euclidean_cluster(p, 0.0), O.euclidean_cluster(*soa(p), 0.0))
    assert R.euclidean_cluster(np.zeros((0, 3), np.float32), 0.1) == []
    one = R.euclidean_cluster(np.array([[1, 2, 3]], np.float32), 0.5)
    assert len(one) == 1 and one[0].tolist() == [0]


def test_labels_are_smallest_index_and_chains_merge():
    """A chain of points 0.9 tol apart in shuffled order forms one cluster (unions across many hops)."""
    import ctypes as C

    from rvpoint_amd._lib import call
    n = 5000
    rng = np.random.default_rng(8)
    t = np.arange(n, dtype=np.float64) * 0.009
    p = np.stack([np.cos(t) * (1 + t), np.sin(t) * (1 + t), t * 0.01], 1).astype(np.float32)
    perm = rng.permutation(n)
    p = p[perm]
    got = R.euclidean_cluster(p, 0.012)
    assert same(got, O.euclidean_cluster(*soa(p), 0.012))
    h = R.SpatialHash()
    h.setInputCloud(p, 0.012)
    h.build()
    lab = R.DeviceArray(n, np.int32)
    idx = R.DeviceArray(n, np.int32)
    off = R.DeviceArray(n + 1, np.int64)
    m = C.c_size_t()
    call("rvp_euclidean_cluster", h.handle, 0.012, 1, 1 << 30, idx.ptr, off.ptr, C.byref(m), lab.ptr, None)
    L = lab.numpy()
    for c in got:
        assert (L[c] == c.min()).all()
    errs = C.c_ulonglong()
    call("rvp_cluster_errors", C.byref(errs))
    assert errs.value == 0
