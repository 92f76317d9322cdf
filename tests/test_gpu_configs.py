"""The benchmark workloads themselves against the oracle, at their real sizes where the oracle finishes in seconds.

c1: the 1.04M-point LiDAR frame through the device frame pipeline (voxel 0.1 -> exact-kNN SOR k=50 -> radius
    normals r=0.5 capped at 64 -> RANSAC 1000, thr 0.05), every stage against the oracle: voxel centroids, SOR
    means and kept points, RANSAC count / model / iterations bit-exact; normals against the FP64 oracle on the
    same (d2, idx) neighbourhoods.  The stage chain is the reference's pipeline driver
    (src/tools/pipeline_export.cpp:356-480).
c3: a 24-cloud subset of the 256-cloud batch (clouds 0..37 of at most 25k points) through the batched stages at
    default settings and with runs of 16 and 64 consecutive SOR queries per wave (cloud boundaries inside runs).
c0: the 100k-point plane+sphere cloud's k=24 normals: the precision figures of that config.
Every test prints its precision figures (median / p99 / max angle) as a JSON line."""
import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes
from rvpoint_amd.pipeline import DeviceFramePipeline, FrameStages

from precision import angles_deg, report

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


def soa(p):
    return [np.ascontiguousarray(p[:, i]) for i in range(3)]


def voxel_centroid_error(p, leaf, got):
    """Voxel centroids (fp32, the reference's sequential sums) against the FP64 mean of each voxel's points:
    voxels keyed as voxel_grid_downsamp_sc keys them (floor(fl(x * fl(1/leaf))), voxel_grid_downsamp.cpp:189-194)
    and ordered lexicographically; returns the median and max distance in metres."""
    inv = np.float32(1.0) / np.float32(leaf)
    keys = np.floor(p * inv).astype(np.int64)
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    k = keys[order]
    head = np.r_[True, (k[1:] != k[:-1]).any(1)]
    gid = np.cumsum(head) - 1
    cnt = np.bincount(gid)
    mean = np.stack([np.bincount(gid, p[order, i].astype(np.float64)) for i in range(3)], 1) / cnt[:, None]
    assert len(mean) == len(got)
    err = np.linalg.norm(got.astype(np.float64) - mean, axis=1)
    return {"median": float(np.median(err)), "max": float(err.max())}


def test_lidar_frame_pipeline_vs_oracle():
    frame = scenes.lidar_frame()
    assert 1_000_000 < len(frame) < 1_100_000
    st = FrameStages()
    pipe = DeviceFramePipeline(frame, st, keep_means=True)
    pipe.step()
    res = pipe.results()
    vox = O.voxel_sc(frame, st.leaf)
    assert res["voxels"].shape == vox.shape and (res["voxels"] == vox).all()
    kept, means = O.sor_sc_mt(vox, st.sor_k, st.sor_alpha)
    assert (res["means"] == means).all()
    assert res["kept"].shape == kept.shape and (res["kept"] == kept).all()
    x, y, z = soa(kept)
    best, model, its = O.ransac_plane(x, y, z, st.thr, st.iters, st.collinear, st.probability)
    assert res["inliers"] == best and (res["model"] == model).all() and res["iterations"] == its
    ref = O.normals_radius_f64(x, y, z, st.radius, st.radius, st.max_nn, st.vp)
    zero_ref = np.linalg.norm(ref, axis=1) == 0
    got = res["normals"]
    assert ((np.linalg.norm(got, axis=1) == 0) == zero_ref).all()
    ang = angles_deg(got[~zero_ref], ref[~zero_ref])
    # the flip rule (normal_estimation.cpp:55-67) toward the sensor at (0, 0, 1.7), evaluated in fp64
    v = np.asarray(st.vp, np.float64) - kept.astype(np.float64)
    assert ((v * got.astype(np.float64)).sum(1)[~zero_ref] >= -1e-4).all()
    rep = report("c1 lidar frame: radius 0.5 cap 64 normals of the SOR output", ang,
                 {"voxels": len(vox), "kept": len(kept), "zero_normals": int(zero_ref.sum()),
                  "fp64_fallbacks": R.normals_fallback_count(),
                  "voxel_centroid_err_m": voxel_centroid_error(frame, st.leaf, res["voxels"])})
    assert rep["angle_deg_median"] <= 0.01 and rep["angle_deg_p99"] <= 1.0


C3_SUBSET = [0, 1, 2, 3, 4, 5, 9, 10, 12, 13, 14, 16, 21, 22, 23, 24, 25, 27, 29, 30, 31, 35, 36, 37]


@pytest.fixture(scope="module")
def c3_subset():
    clouds, scales = scenes.batch_clouds(with_scales=True, which=C3_SUBSET)
    assert all(len(c) <= 25_000 for c in clouds)
    return clouds, scales


@pytest.mark.parametrize("per_wave,steal,start", [(0, 0, 1), (0, 1, 1), (0, 2, 1), (16, 1, 1), (64, 2, 1),
                                                  (0, 0, 3), (4, 2, 2)])
def test_batch_subset_stages_vs_oracle(c3_subset, per_wave, steal, start):
    """SOR k=32 -> radius r=0.02*scale capped at 64 on the kept points -> kNN normals k=24, as the batched leg
    of the benchmark; per_wave 0 = the run length chosen from the count, else forced; steal: persistent waves with
    XCD-local runs and stealing (1), runs strided within per-XCD ranges (2), or over the launch (0); start: the
    half-width in cells of an unbounded k-NN search's first box."""
    clouds, scales = c3_subset
    b = R.Batch(clouds)
    knobs = ("queries_per_wave", "work_steal", "knn_start_cells")
    old = [R._lib.tune(n) for n in knobs]
    try:
        for n, v in zip(knobs, (per_wave, steal, start)):
            R._lib.tune(n, v)
        kept, means = R.batch_sor_exact(b, 32, 1.0)
        kept2, means2 = R.batch_sor_exact(b, 32, 1.0)  # the work counters start from zero again
        nrm_start = R.batch_normals_knn(R.Batch(kept), 24) if start != 1 else None
    finally:
        for n, v in zip(knobs, old):
            R._lib.tune(n, v)
    assert (means2 == means).all() and all((a == c).all() for a, c in zip(kept, kept2))
    okept = []
    for c, p in enumerate(clouds):
        oout, omeans = O.sor_sc_mt(p, 32, 1.0)
        assert (b.split(means)[c] == omeans).all(), c
        assert kept[c].shape == oout.shape and (kept[c] == oout).all(), c
        okept.append(oout)
    if start != 1:
        # the k-NN normals of the same kept points with the default first box: the same neighbourhoods
        assert all((a == c).all() for a, c in zip(R.Batch(kept).split(nrm_start),
                                                  R.Batch(kept).split(R.batch_normals_knn(R.Batch(kept), 24))))
    if per_wave or steal or start != 1:
        return
    kb = R.Batch(okept)
    radius = np.array([0.02 * s for s in scales], np.float32)
    idx, d2, cnt = R.batch_radius_search(kb, radius, 64)
    nrm = R.batch_normals_knn(kb, 24)
    angs = []
    for c, p in enumerate(okept):
        x, y, z = soa(p)
        oi, od, oc, _ = O.hash_radius(x, y, z, float(radius[c]), x, y, z, float(radius[c]), 64)
        ci, cd, cc = kb.split(idx)[c], kb.split(d2)[c], kb.split(cnt)[c]
        assert (cc == oc).all(), c
        for i in range(len(p)):
            m = oc[i]
            assert (ci[i, :m] == oi[i, :m]).all() and (cd[i, :m] == od[i, :m]).all(), (c, i)
        angs.append(angles_deg(kb.split(nrm)[c], O.normals_knn_f64(x, y, z, 24)))
    rep = report("c3 subset: k=24 normals of the SOR output, 24 clouds", np.concatenate(angs))
    assert rep["angle_deg_median"] <= 0.01 and rep["angle_deg_p99"] <= 1.0


@pytest.mark.parametrize("per_wave,start", [(16, 1), (64, 1), (2, 3), (0, 8)])
def test_single_cloud_sor_runs_of_queries(per_wave, start):
    """Single-cloud exact SOR with forced runs of consecutive queries per wave and first k-NN boxes of
    half-width `start` cells."""
    p = scenes.plane_sphere_cloud(6000, seed=77)
    old = R._lib.tune("queries_per_wave"), R._lib.tune("knn_start_cells")
    try:
        R._lib.tune("queries_per_wave", per_wave)
        R._lib.tune("knn_start_cells", start)
        out, means = R.sor_sc(p, 30, 1.0, with_means=True)
    finally:
        R._lib.tune("queries_per_wave", old[0])
        R._lib.tune("knn_start_cells", old[1])
    oout, omeans = O.sor_sc(p, 30, 1.0)
    assert (means == omeans).all()
    assert out.shape == oout.shape and (out == oout).all()


def test_c0_normals_precision():
    """c0 (100k plane+sphere points, 2% outliers): k=24 normals against the FP64 oracle on a 20k-point sample of
    the same neighbourhoods (the k+1 nearest by (d2, idx))."""
    p = scenes.plane_sphere_cloud()
    x, y, z = soa(p)
    nx, ny, nz = R.normal_estimation_sc(p, 24, 0.0)
    fb = R.normals_fallback_count()
    sel = np.sort(np.random.default_rng(42).choice(len(p), 20_000, replace=False))
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.02)
    h.build()
    idx, _ = h.knn(p[sel], 25)
    ref = np.empty((len(sel), 3))
    for j in range(len(sel)):
        ref[j], _ = O.normal_from_points_f64(p[idx[j]].astype(np.float64))
    ang = angles_deg(np.c_[nx, ny, nz][sel], ref)
    rep = report("c0 plane+sphere 100k: k=24 normals", ang, {"fp64_fallback_fraction": fb / len(p)})
    assert rep["angle_deg_median"] <= 0.01 and rep["angle_deg_p99"] <= 1.0
