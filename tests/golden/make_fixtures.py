"""Writes the committed fixtures under tests/golden/ (data only: inputs and expected outputs).

kitti_0000000000.npz : xyz (float32, n x 3) of the reference's data/0000000000.pcd (PCD v0.7, binary,
                       FIELDS x y z), read with the small PCD reader below.  The reference's documented known
                       answers on this cloud (docs/PIPELINE_PROFILING_AND_ABLATION_EXPERIMENT.md:52-53,121,123)
                       are checked by tests/test_oracle_kat.py.
Run in a container where the reference tree is mounted:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def read_pcd_xyz(path):
    """Minimal PCD v0.7 reader for x/y/z float fields, DATA ascii or binary."""
    with open(path, "rb") as f:
        header = {}
        while True:
            line = f.readline().decode("ascii", "replace").strip()
            if not line or line.startswith("#"):
                continue
            key, _, val = line.partition(" ")
            header[key.upper()] = val.split()
            if key.upper() == "DATA":
                break
        fields = header["FIELDS"]
        sizes = [int(s) for s in header["SIZE"]]
        types = header["TYPE"]
        counts = [int(c) for c in header.get("COUNT", ["1"] * len(fields))]
        npts = int(header["POINTS"][0])
        kind = header["DATA"][0]
        if kind == "ascii":
            rows = np.loadtxt(f, dtype=np.float64, ndmin=2)[:npts]
            cols = {fld: rows[:, sum(counts[:i])] for i, fld in enumerate(fields)}
        elif kind == "binary":
            dt = np.dtype([(fld, {("F", 4): "<f4", ("F", 8): "<f8", ("U", 4): "<u4", ("I", 4): "<i4",
                                  ("U", 1): "u1", ("I", 1): "i1", ("U", 2): "<u2", ("I", 2): "<i2"}[(t, s)], (c,))
                           for fld, s, t, c in zip(fields, sizes, types, counts)])
            rec = np.frombuffer(f.read(dt.itemsize * npts), dtype=dt, count=npts)
            cols = {fld: rec[fld][:, 0] for fld in fields}
        else:
            raise ValueError(f"unsupported PCD DATA kind {kind}")
    return np.stack([cols["x"], cols["y"], cols["z"]], 1).astype(np.float32)


def main():
    src = os.path.join(REF, "data", "0000000000.pcd")
    if not os.path.exists(src):
        sys.exit(f"{src} not found (run where the reference tree is mounted)")
    xyz = read_pcd_xyz(src)
    np.savez_compressed(os.path.join(HERE, "kitti_0000000000.npz"), xyz=xyz)
    print("wrote kitti_0000000000.npz", xyz.shape)


if __name__ == "__main__":
    main()
