"""PCD ingest on the host: the reference's own data files against the reference's loader (oracle/_ref), the
documented point count of 0000000000.pcd and the committed fixture, the writer against savePCD, and
synthetic ascii / binary / binary_compressed files with unaligned record layouts."""
import os

import numpy as np
import pytest

import oracle as O
from pcd_util import lzf_encode, write_pcd

import rvpoint_amd as R

DATA = "/root/reference/data"
needs_ref = pytest.mark.skipif(not (O.ref_available() and os.path.isdir(DATA)), reason="reference tree absent")
HERE = os.path.dirname(os.path.abspath(__file__))


@needs_ref
@pytest.mark.parametrize("name", ["0000000000.pcd", "0000000010.pcd", "bunny.pcd", "indoor_scene.pcd",
                                  "living_room.pcd", "table_scene_lms400.pcd"])
def test_reference_data_files_equal_reference_loader(name):
    path = os.path.join(DATA, name)
    n, mode = R.pcd_info(path)
    got = R.load_pcd(path)
    ref = O.ref_load_pcd(path, max(n, len(got)) + 16)
    assert ref is not None and got.shape == ref.shape and np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@needs_ref
def test_kitti_point_count_and_fixture():
    path = os.path.join(DATA, "0000000000.pcd")
    assert R.pcd_info(path) == (114278, "binary")
    fx = np.load(os.path.join(HERE, "golden", "kitti_0000000000.npz"))["xyz"]
    assert np.array_equal(R.load_pcd(path), fx)


@needs_ref
@pytest.mark.parametrize("binary", [False, True])
def test_writer_equals_savepcd(tmp_path, binary):
    p = (np.random.default_rng(3).normal(0, 10, (500, 3))).astype(np.float32)
    a, b = tmp_path / "ours.pcd", tmp_path / "ref.pcd"
    R.save_pcd(a, p, binary)
    O.ref_save_pcd(b, p, binary)
    assert a.read_bytes() == b.read_bytes()


def test_lzf_encoder_round_trip():
    rng = np.random.default_rng(1)
    for data in [b"", b"a", bytes(1000), rng.integers(0, 4, 5000).astype(np.uint8).tobytes(),
                 rng.integers(0, 256, 3000).astype(np.uint8).tobytes()]:
        enc = lzf_encode(data)
        assert _lzf_decode(enc, len(data)) == data


def _lzf_decode(buf, n):
    out = bytearray()
    i = 0
    while i < len(buf):
        c = buf[i]
        i += 1
        if c < 32:
            out += buf[i:i + c + 1]
            i += c + 1
        else:
            L = c >> 5
            if L == 7:
                L += buf[i]
                i += 1
            d = ((c & 31) << 8) + buf[i] + 1
            i += 1
            for _ in range(L + 2):
                out.append(out[-d])
    assert len(out) == n
    return bytes(out)


@pytest.mark.parametrize("mode", ["ascii", "binary", "binary_compressed"])
def test_synthetic_layouts(tmp_path, mode):
    rng = np.random.default_rng(11)
    n = 3000
    xyz = rng.normal(0, 5, (n, 3)).astype(np.float32)
    xyz[::7] = np.round(xyz[::7])                      # repeats for the compressor
    cols = [("intensity", rng.integers(0, 255, n).astype(np.uint8), "U"), ("x", xyz[:, 0], "F"),
            ("ring", rng.integers(0, 64, n).astype(np.uint16), "U"), ("y", xyz[:, 1], "F"), ("z", xyz[:, 2], "F")]
    path = tmp_path / f"t_{mode}.pcd"
    write_pcd(path, cols, mode)
    assert R.pcd_info(path) == (n, mode)
    got = R.load_pcd(path)
    if mode == "ascii":
        assert np.array_equal(got, np.array([[float(repr(float(v))) for v in r] for r in xyz], np.float32))
    else:
        assert np.array_equal(got, xyz)


def test_errors(tmp_path):
    with pytest.raises(R.RvpError):
        R.load_pcd(tmp_path / "missing.pcd")
    bad = tmp_path / "bad.pcd"
    bad.write_bytes(b"VERSION 0.7\nFIELDS a b\nSIZE 4 4\nTYPE F F\nPOINTS 3\nDATA binary\n" + bytes(24))
    with pytest.raises(R.RvpError):
        R.load_pcd(bad)
