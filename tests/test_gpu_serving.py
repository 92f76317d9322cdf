"""Streaming frames with continuous batching: every frame's results in a mixed batch equal the frame run
alone, the voxel -> SOR -> RANSAC chain equals the oracle's, and the served normals meet the FP64 oracle's
bar (the same radius neighbourhoods, capped as the server caps them)."""
import numpy as np
import pytest

import oracle as O
from rvpoint_amd import scenes

from precision import angles_deg, report

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


@pytest.fixture(scope="module")
def frames():
    # small LiDAR-like frames of uneven sizes (rings x azimuths), as the streaming workload's generator
    shapes = [(8, 1500, 5001), (16, 2500, 5002), (4, 900, 5003), (32, 1200, 5004), (12, 3000, 5005)]
    return [scenes.lidar_frame(rings=r, azimuths=a, seed=sd, nbox=12, max_range=40.0) for r, a, sd in shapes]


def params():
    from rvpoint_amd.serve import FrameParams
    import ctypes as C
    return FrameParams(0.1, 20, 1.0, 0.5, 32, (C.c_float * 3)(0.0, 0.0, 1.7), 0.05, 200, 0.99)


def same(a, b):
    return (a["n_voxel"] == b["n_voxel"] and a["n_sor"] == b["n_sor"] and np.array_equal(a["points"], b["points"])
            and np.array_equal(a["normals"], b["normals"]) and np.array_equal(a["model"], b["model"])
            and a["inliers"] == b["inliers"] and a["iterations"] == b["iterations"])


def test_frame_results_same_alone_and_in_the_mix(frames):
    from rvpoint_amd.serve import FrameServer
    srv = FrameServer(params(), slots=2, slot_points=200_000, hbm_budget=4 << 30)
    try:
        alone = [srv.wait(srv.submit(f)) for f in frames]           # one frame per batch
        assert all(r["batch_frames"] == 1 for r in alone)
        tickets = [srv.submit(f) for f in frames + frames[::-1]]     # queued together: batched, across slots
        mixed = [srv.wait(t) for t in tickets]
        assert max(r["batch_frames"] for r in mixed) > 1
        for i, r in enumerate(mixed):
            j = i if i < len(frames) else len(frames) - 1 - (i - len(frames))
            assert same(r, alone[j]), i
    finally:
        srv.close()


def test_served_chain_matches_oracle(frames):
    from rvpoint_amd.serve import FrameServer
    srv = FrameServer(params(), slots=1, slot_points=200_000, hbm_budget=4 << 30)
    try:
        for f in frames[:3]:
            r = srv.wait(srv.submit(f))
            vox = O.voxel_sc(f, 0.1)
            kept, _ = O.sor_sc(vox, 20, 1.0)
            assert r["n_in"] == len(f) and r["n_voxel"] == len(vox)
            assert np.array_equal(r["points"], kept)
            best, model, its = O.ransac_plane(*(np.ascontiguousarray(kept[:, i]) for i in range(3)), 0.05, 200)
            assert r["inliers"] == best and np.array_equal(r["model"], model) and r["iterations"] == its
            assert np.allclose(np.linalg.norm(r["normals"][r["normals"].any(1)], axis=1), 1.0, atol=1e-5)
    finally:
        srv.close()


def test_served_normals_vs_fp64(frames):
    """Two slots on two streams with frames queued together: each frame's normals against the FP64 oracle's
    radius normals (r = 0.5, cap 32, viewpoint (0, 0, 1.7)) of its kept points."""
    from rvpoint_amd.serve import FrameServer
    P = params()
    srv = FrameServer(P, slots=2, slot_points=200_000, hbm_budget=4 << 30)
    try:
        res = [srv.wait(t) for t in [srv.submit(f) for f in frames + frames[::-1]]]
    finally:
        srv.close()
    angs = []
    for r in res:
        kept = r["points"]
        x, y, z = (np.ascontiguousarray(kept[:, i]) for i in range(3))
        ref = O.normals_radius_f64(x, y, z, P.normal_radius, P.normal_radius, P.normal_max_nn,
                                   (P.vp[0], P.vp[1], P.vp[2]))
        zero_ref = np.linalg.norm(ref, axis=1) == 0
        got = r["normals"]
        assert ((np.linalg.norm(got, axis=1) == 0) == zero_ref).all()
        angs.append(angles_deg(got[~zero_ref], ref[~zero_ref]))
    rep = report("served frames (2 slots): radius 0.5 cap 32 normals", np.concatenate(angs))
    assert rep["angle_deg_median"] <= 0.01 and rep["angle_deg_p99"] <= 1.0


def test_server_rejects_oversized_frames(frames):
    from rvpoint_amd.serve import FrameServer
    srv = FrameServer(params(), slots=1, slot_points=4096, hbm_budget=1 << 30)
    try:
        with pytest.raises(R.RvpError):
            srv.submit(frames[1])
    finally:
        srv.close()
