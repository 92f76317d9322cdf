"""The device frame pipeline replayed as a captured HIP graph: every replay equals the eager pipeline's results
(voxel centroids, SOR means and kept points, normals, RANSAC plane), with eager work and synchronisations of the
program between replays, and a new frame loaded into the same buffers between replays is processed from its own
contents (the graph reads device-side counts, and its look-back epochs advance on the device)."""
import numpy as np
import pytest

from rvpoint_amd import scenes
from rvpoint_amd.pipeline import ConcurrentFrames, DeviceFramePipeline, PipelinedFrames

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("voxels", "means", "kept", "normals", "model")) and \
        a["inliers"] == b["inliers"] and a["iterations"] == b["iterations"]


def eager(frame):
    p = DeviceFramePipeline(frame, keep_means=True)
    p.step()
    return p.results()


def test_frame_graph_replays_equal_eager():
    f1 = scenes.lidar_frame(rings=24, azimuths=6000, seed=11, nbox=14, max_range=40.0)
    f2 = scenes.lidar_frame(rings=12, azimuths=5000, seed=12, nbox=10, max_range=40.0)
    assert len(f2) < len(f1)
    e1, e2 = eager(f1), eager(f2)
    g = DeviceFramePipeline(f1, keep_means=True, graph=True)
    try:
        for rep in range(3):
            g.step()
            assert g.graph is not None
            assert same(g.results(), e1), rep
            # the program's own eager work between replays, on the null stream, then a full synchronisation
            R.sor_sc(f2[:3000], 20, 1.0)
            R.synchronize()
        g.load(f2)
        g.step()
        assert same(g.results(), e2)
        g.load(f1)
        g.step()
        g.step()
        assert same(g.results(), e1)
    finally:
        g.close()
    import ctypes as C
    errs = C.c_ulonglong()
    R._lib.call("rvp_scan_wait_errors", C.byref(errs))
    assert errs.value == 0


def test_pipelined_frames_equal_eager():
    """Double-buffered frames: frame i + 1's graph is enqueued before frame i's RANSAC (on a second stream) and
    its host replay; every frame's results equal the eager pipeline's, with frames of different sizes in turn."""
    f1 = scenes.lidar_frame(rings=24, azimuths=6000, seed=11, nbox=14, max_range=40.0)
    f2 = scenes.lidar_frame(rings=12, azimuths=5000, seed=12, nbox=10, max_range=40.0)
    f3 = scenes.lidar_frame(rings=16, azimuths=4000, seed=13, nbox=8, max_range=30.0)
    frames = [f1, f2, f3, f2, f1, f3, f3]
    ref = [eager(f) for f in (f1, f2, f3)]
    which = [0, 1, 2, 1, 0, 2, 2]
    pp = PipelinedFrames(f1, keep_means=True, n_max=max(len(f) for f in (f1, f2, f3)))
    got = []
    try:
        for i, f in enumerate(frames):
            pp.load(f)
            pp.step()
            if i >= 1:  # step i finished frame i - 1, held by pipeline (i - 1) % 2 until step i + 1
                assert pp.finished[-1] == (i - 1, (i - 1) % 2)
                got.append(pp.pipes[(i - 1) % 2].results())
        pp.drain()
        got.append(pp.pipes[(len(frames) - 1) % 2].results())
    finally:
        pp.close()
    assert len(got) == len(frames)
    for i, (w, r) in enumerate(zip(which, got)):
        assert same(r, ref[w]), i


def test_concurrent_frames_equal_eager():
    """Three pipelines on their own streams and host threads, each replaying its graph over a different frame
    while the others run: every frame's results equal the eager pipeline's, checked after each frame."""
    fs = [scenes.lidar_frame(rings=24, azimuths=6000, seed=11, nbox=14, max_range=40.0),
          scenes.lidar_frame(rings=12, azimuths=5000, seed=12, nbox=10, max_range=40.0),
          scenes.lidar_frame(rings=16, azimuths=4000, seed=13, nbox=8, max_range=30.0)]
    ref = [eager(f) for f in fs]
    cf = ConcurrentFrames(fs[0], n=3, keep_means=True, n_max=max(len(f) for f in fs))
    bad = []
    try:
        for i in range(3):
            cf.load(i, fs[i])

        def check(i, j):
            if not same(cf.pipes[i].results(), ref[i]):
                bad.append((i, j))

        wall, lat = cf.run(4, on_frame=check)
    finally:
        cf.close()
    assert len(lat) == 12 and not bad, bad
