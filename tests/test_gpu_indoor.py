"""Dense indoor scan, 20M points (10 x 8 x 3 m room, 60 boxes, 20 cylinders, sigma 3 mm):
FP32-compensated normals (k=30) and neighbourhood centroids against the FP64 oracle on the same
neighbourhoods, and spatial-hash radius lists (r = 0.02, cap 64) against the oracle's hash search on a
sample of queries.  Prints the precision report (median / p99 angle and centroid error, FP64 fallbacks)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes

from precision import report

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

R = pytest.importorskip("rvpoint_amd")
K = 30
SAMPLE = 20_000


@pytest.fixture(scope="module")
def indoor():
    pts = scenes.indoor_scan()
    return pts, R.DeviceCloud(pts)


def test_indoor_normals_precision_vs_fp64(indoor):
    from rvpoint_amd._lib import call
    pts, dev = indoor
    n = len(pts)
    D = R.DeviceArray
    nx, ny, nz, cx, cy, cz = (D(n) for _ in range(6))
    call("rvp_normals_knn_centroids", *dev.ptrs(), n, K, 0.0, 0.0, 0.0, nx.ptr, ny.ptr, nz.ptr, cx.ptr, cy.ptr,
         cz.ptr, None)
    fb = R.normals_fallback_count()
    rng = np.random.default_rng(2002)
    sel = np.sort(rng.choice(n, SAMPLE, replace=False))
    nrm = np.stack([nx.numpy()[sel], ny.numpy()[sel], nz.numpy()[sel]], 1).astype(np.float64)
    cen = np.stack([cx.numpy()[sel], cy.numpy()[sel], cz.numpy()[sel]], 1).astype(np.float64)
    # the same neighbourhoods (k+1 nearest by (d2, idx), self included) from the k-NN query
    h = R.SpatialHash.from_device(dev, 0.02)
    idx, _ = h.knn(pts[sel], K + 1)
    ang = np.empty(SAMPLE)
    cerr = np.empty(SAMPLE)
    for j in range(SAMPLE):
        P = pts[idx[j]].astype(np.float64)
        n64, _ = O.normal_from_points_f64(P)
        ang[j] = np.degrees(np.arccos(min(1.0, abs(float(nrm[j] @ n64)))))
        cerr[j] = np.linalg.norm(cen[j] - P.mean(0))
    rep = report("c2 indoor 20M: k=30 normals and neighbourhood centroids (20k-point sample)", ang,
                 {"cloud_points": n, "k": K, "centroid_err_m_median": float(np.median(cerr)),
                  "centroid_err_m_p99": float(np.percentile(cerr, 99)), "centroid_err_m_max": float(cerr.max()),
                  "fp64_fallback_fraction": fb / n})
    assert rep["angle_deg_median"] <= 0.01
    assert rep["angle_deg_p99"] <= 1.0
    assert rep["centroid_err_m_max"] <= 1e-5


def test_indoor_radius_lists_vs_oracle_sample(indoor):
    pts, dev = indoor
    rng = np.random.default_rng(2003)
    q = pts[rng.choice(len(pts), 2000, replace=False)]
    h = R.SpatialHash.from_device(dev, 0.02)
    got = h.radius_search_batch(q, 0.02, 64)
    x, y, z = (np.ascontiguousarray(pts[:, i]) for i in range(3))
    qx, qy, qz = (np.ascontiguousarray(q[:, i]) for i in range(3))
    oi, od, oc, _ = O.hash_radius(x, y, z, 0.02, qx, qy, qz, 0.02, 64)
    for j, (gi, gd) in enumerate(got):
        m = oc[j]
        assert len(gi) == m and (gi == oi[j, :m]).all() and (gd == od[j, :m]).all(), j
