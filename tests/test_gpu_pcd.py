"""PCD ingest straight into device memory: the device path (binary de-interleave on the GPU, compressed
blocks decoded on the host) gives the same bits as the host reader, for unaligned record layouts, every
data mode, and the KITTI frame fixture written back out as binary."""
import os

import numpy as np
import pytest

from pcd_util import write_pcd

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")
HERE = os.path.dirname(os.path.abspath(__file__))


def device_xyz(path):
    d = R.load_pcd_device(path)
    return np.stack([d.x.numpy()[:d.n], d.y.numpy()[:d.n], d.z.numpy()[:d.n]], 1)


@pytest.mark.parametrize("mode", ["ascii", "binary", "binary_compressed"])
@pytest.mark.parametrize("n", [1, 999, 70001])
def test_device_reader_equals_host_reader(tmp_path, mode, n):
    if mode == "ascii" and n > 1000:
        pytest.skip("ascii writer is slow in Python at this size")
    rng = np.random.default_rng(n)
    xyz = rng.normal(0, 30, (n, 3)).astype(np.float32)
    cols = [("t", rng.integers(0, 1 << 30, n).astype(np.uint32), "U"), ("x", xyz[:, 0], "F"),
            ("ring", rng.integers(0, 64, n).astype(np.uint8), "U"), ("y", xyz[:, 1], "F"),
            ("z", xyz[:, 2], "F"), ("i", rng.normal(0, 1, n).astype(np.float64), "F")]
    path = tmp_path / f"c_{mode}_{n}.pcd"
    write_pcd(path, cols, mode)
    host = R.load_pcd(path)
    dev = device_xyz(path)
    assert np.array_equal(host.view(np.uint32), dev.view(np.uint32))
    if mode != "ascii":
        assert np.array_equal(dev, xyz)


def test_kitti_fixture_round_trip(tmp_path):
    xyz = np.load(os.path.join(HERE, "golden", "kitti_0000000000.npz"))["xyz"]
    path = tmp_path / "kitti.pcd"
    R.save_pcd(path, xyz, binary=True)
    assert R.pcd_info(path) == (len(xyz), "binary")
    assert np.array_equal(device_xyz(path), xyz)
