"""Parity of the smaller pieces of the path against the oracle:

- D1 / D2 through the C ABI (rvp_dist_sq, rvp_inds_in_radius) and through the C++ drop-in wrappers
  (get_dist_sq_rvv, get_inds_in_radius_rvv: bits, subset order, appending without clearing;
  src/core/rvv_common.cpp:36-68);
- the FP64 eigen fallback on near-degenerate neighbourhoods (exact cigar-shaped and collinear clusters),
  counted and graded against the FP64 oracle;
- the reference's documented SOR known answers through the HIP path on data/0000000000.pcd voxelised at 0.20
  (docs/PIPELINE_PROFILING_AND_ABLATION_EXPERIMENT.md:121,123: sor_spatial_hash r=0.5 keeps 16,613, sor_sc
  k=20 keeps 17,569);
- look-back primitives on freshly created non-blocking streams (their state is cleared in stream order)."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes

from precision import angles_deg, report

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def soa(p):
    return [np.ascontiguousarray(p[:, i]) for i in range(3)]


# ------------------------------------------------------------------------------------------------ D1 / D2
def test_dist_sq_c_abi_bits():
    from rvpoint_amd._lib import call
    rng = np.random.default_rng(51)
    p = np.concatenate([rng.normal(0, 3, (20000, 3)), rng.uniform(-1e4, 1e4, (1000, 3))]).astype(np.float32)
    dev = R.DeviceCloud(p)
    out = R.DeviceArray(len(p))
    for q in (np.float32([0.1, -0.2, 0.3]), np.float32([1e4, -5.5, 123.25])):
        call("rvp_dist_sq", *dev.ptrs(), len(p), *[float(v) for v in q], out.ptr, None)
        assert (out.numpy() == O.dist_sq(*soa(p), q)).all()


@pytest.mark.parametrize("r2", [0.0, 0.5, 4.0, 1e9])
def test_inds_in_radius_c_abi_subset_order(r2):
    from rvpoint_amd._lib import call
    rng = np.random.default_rng(52)
    p = rng.normal(0, 1, (30000, 3)).astype(np.float32)
    sub = rng.permutation(len(p))[:20000].astype(np.int32)
    sub[:50] = sub[50:100]  # repeated subset entries are reported each time, in subset order
    dev = R.DeviceCloud(p)
    dsub = R.DeviceArray.from_numpy(sub)
    oi, od = R.DeviceArray(len(sub), np.int32), R.DeviceArray(len(sub))
    q = np.float32([0.25, 0.5, -0.125])
    m = C.c_size_t()
    call("rvp_inds_in_radius", *dev.ptrs(), dsub.ptr, len(sub), *[float(v) for v in q], float(np.float32(r2)),
         oi.ptr, od.ptr, C.byref(m), None)
    ei, ed = O.inds_in_radius(*soa(p), sub, q, np.float32(r2))
    assert m.value == len(ei)
    assert (oi.numpy()[:m.value] == ei).all() and (od.numpy()[:m.value] == ed).all()


def test_cpp_distance_wrappers_vs_oracle():
    """get_dist_sq_rvv / get_inds_in_radius_rvv of the C++ drop-in: the program checks appending itself and
    dumps its inputs and outputs, which are compared here with the oracle."""
    exe = os.path.join(ROOT, "tests", "cpp", "_build", "unit_checks")
    assert os.path.exists(exe), "tests/cpp/_build/unit_checks not built (python -m rvpoint_amd.build)"
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "dist.bin")
        r = subprocess.run([exe, "--dump", path], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        raw = open(path, "rb").read()
    n, m, k = np.frombuffer(raw[:24], np.uint64).astype(int)
    q = np.frombuffer(raw[24:40], np.float32)
    off = 40
    arrs = []
    for cnt, dt in ((n, np.float32), (n, np.float32), (n, np.float32), (m, np.int32), (n, np.float32),
                    (k, np.int32), (k, np.float32)):
        arrs.append(np.frombuffer(raw[off:off + 4 * cnt], dt))
        off += 4 * cnt
    x, y, z, sub, d, idx, dd = arrs
    assert (d == O.dist_sq(x, y, z, q[:3])).all()
    ei, ed = O.inds_in_radius(x, y, z, sub, q[:3], q[3], [-1, -2, -3], [9.0, 8.0, 7.0])
    assert len(ei) == k and (idx == ei).all() and (dd == ed).all()


# ------------------------------------------------------------------------------------------------ FP64 fallback
def degenerate_clusters():
    """Clusters of exactly 32 points (k = 31 gives each point its own cluster as neighbourhood) on a 2^-8 grid,
    so that every difference, product and sum of the covariance is exact in fp32:
    - cigars: offsets +-(512,0,0), +-(0,16,16), +-(0,16,-16), +-(0,a,a) and 24 points at the centre, under a
      signed axis permutation (in units of 2^-8): the covariance has eigenvalues 2*512^2, 1024 + 4a^2 and 1024,
      the two smallest 4a^2 apart (a = 1, 2): relative gaps of 7.6e-6 and 3.1e-5, normal along (0, 1, -1);
    - collinear strands: t * (1, 2, 3) for t = -15 .. 16 (rank-one covariance, no defined normal)."""
    perms = [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0), (1, 0, 2)]
    rng = np.random.default_rng(60)
    pts, kinds = [], []
    c = 0
    for a in (1, 2):
        base = np.array([[512, 0, 0], [0, 16, 16], [0, 16, -16], [0, a, a]], np.float64)
        off = np.concatenate([base, -base, np.zeros((24, 3))])
        for pm in perms:
            for sg in range(4):
                signs = np.array([1 if (sg >> b) & 1 == 0 else -1 for b in range(3)], np.float64)
                o = off[:, pm] * signs
                centre = np.array([c % 8, (c // 8) % 8, c // 64], np.float64) * 16.0
                pts.append((centre + o / 256.0))
                kinds.append(np.full(32, 0))
                c += 1
    for j in range(12):
        t = np.arange(-15, 17, dtype=np.float64)[:, None]
        d = np.array([1, 2, 3], np.float64)[list(perms[j % 6])]
        centre = np.array([c % 8, (c // 8) % 8, c // 64], np.float64) * 16.0
        pts.append(centre + t * d / 256.0)
        kinds.append(np.full(32, 1))
        c += 1
    p = np.concatenate(pts).astype(np.float32)
    assert (p.astype(np.float64) == np.concatenate(pts)).all()  # exact on the grid
    order = rng.permutation(len(p))
    return p[order], np.concatenate(kinds)[order]


def test_fp64_fallback_on_near_degenerate_neighbourhoods():
    p, kind = degenerate_clusters()
    x, y, z = soa(p)
    for path in ("single", "batched"):
        if path == "single":
            nx, ny, nz = R.normal_estimation_sc(p, 31, 0.0)
            got = np.c_[nx, ny, nz]
        else:
            b = R.Batch([p, p])  # two clouds: each point's neighbourhood stays within its own copy
            both = b.split(R.batch_normals_knn(b, 31))
            assert (both[0] == both[1]).all()
            got = both[0]
        fb = R.normals_fallback_count()
        assert fb >= (1 if path == "single" else 2) * len(p), (path, fb, len(p))  # every cluster is near-degenerate
        ref = O.normals_knn_f64(x, y, z, 31)
        cig = kind == 0
        ang = angles_deg(got[cig], ref[cig])
        rep = report(f"near-degenerate cigars ({path})", ang, {"fp64_fallbacks": int(fb), "points": int(len(p))})
        assert rep["angle_deg_median"] <= 0.01 and rep["angle_deg_max"] <= 0.01
        # collinear strands: any unit vector perpendicular to the strand is an eigenvector of the (zero) smallest
        # eigenvalue
        line = kind == 1
        g = got[line].astype(np.float64)
        assert np.allclose(np.linalg.norm(g, axis=1), 1.0, atol=1e-5)
        for i in np.nonzero(line)[0][:64]:
            cl = np.linalg.norm(p - p[i], axis=1) < 0.5
            v = p[cl].astype(np.float64)
            u = np.linalg.svd(v - v.mean(0))[2][0]
            assert abs(float(u @ got[i])) <= 1e-6, i


# ------------------------------------------------------------------------------------------------ KATs
@pytest.fixture(scope="module")
def kitti_vox():
    xyz = np.load(os.path.join(os.path.dirname(__file__), "golden", "kitti_0000000000.npz"))["xyz"]
    v = R.voxel_grid_downsamp_sc(xyz, 0.20)
    assert len(v) == 18542
    return v


def test_sor_sc_kat_on_gpu(kitti_vox):
    out = R.sor_sc(kitti_vox, 20, 1.0)
    assert len(out) == 17569
    oout, _ = O.sor_sc_mt(kitti_vox, 20, 1.0)
    assert (out == oout).all()


def test_sor_spatial_hash_kat_on_gpu(kitti_vox):
    x, y, z = soa(kitti_vox)
    h = R.SpatialHash()
    h.setInputCloud(R.PointCloudSoA(x, y, z), 0.5)
    h.build()
    out = R.sor_spatial_hash(R.PointCloudSoA(x, y, z), h, 20, 1.0, 0.5)
    assert len(out) == 16613
    oout, _ = O.sor_radius(x, y, z, 0.5, 20, 1.0, 0.5)
    assert (out == oout).all()


# ------------------------------------------------------------------------------------------------ fresh streams
def test_sorts_and_batches_on_fresh_nonblocking_streams():
    """Look-back state (ticket counter, status words) is created per stream and cleared in that stream's order:
    the first sort, scan and compaction on each new non-blocking stream give the oracle's results."""
    clouds = [scenes.plane_sphere_cloud(n, seed=90 + n) for n in (3000, 1200, 5000)]
    for rep in range(3):
        s = R.Stream()
        try:
            b = R.Batch(clouds, stream=s.handle)
            kept, means = R.batch_sor_exact(b, 16, 1.0, stream=s.handle)
            s.sync()
            for c, p in enumerate(clouds):
                oout, omeans = O.sor_sc(p, 16, 1.0)
                assert (b.split(means)[c] == omeans).all(), (rep, c)
                assert (kept[c] == oout).all(), (rep, c)
        finally:
            s.close()
    errs = C.c_ulonglong()
    R._lib.call("rvp_scan_wait_errors", C.byref(errs))
    assert errs.value == 0
