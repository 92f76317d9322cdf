"""The LiDAR-frame stage chain on the device: voxel grid -> exact-kNN SOR -> radius normals -> RANSAC plane.

The chain follows the reference's pipeline driver (src/tools/pipeline_export.cpp:356-480: voxel_grid_downsamp,
sor, normal_estimation with a radius search, ransac_plane), with a cap on the radius neighbours.  The frame runs
as a batch of one cloud: every stage reads its point count from the previous stage's device offsets, so nothing
waits on the host until RANSAC's adaptive stop is replayed at the end.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import call
from .device import DeviceArray, DeviceCloud


@dataclass
class FrameStages:
    """Stage parameters (reference names): voxel leaf, SOR k and std multiplier, normals radius, neighbour cap
    and viewpoint, RANSAC threshold, hypotheses, collinearity threshold and probability."""
    leaf: float = 0.1
    sor_k: int = 50
    sor_alpha: float = 1.0
    radius: float = 0.5
    max_nn: int = 64
    vp: tuple = (0.0, 0.0, 1.7)
    thr: float = 0.05
    iters: int = 1000
    collinear: float = 1e-6
    probability: float = 0.99


class DeviceFramePipeline:
    """One frame resident in HBM through the four stages, stage outputs kept on the device.

    keep_means: also keep SOR's per-point mean k-NN distances (voxel order) for inspection."""

    def __init__(self, frame_xyz, stages: FrameStages | None = None, stream=None, keep_means=False):
        self.st = stages or FrameStages()
        self.s = stream
        self.n = n = len(frame_xyz)
        D = DeviceArray
        self.frame = DeviceCloud(frame_xyz)
        self.off = D.from_numpy(np.array([0, n], np.uint32))
        self.leaf = D.from_numpy(np.array([self.st.leaf], np.float32))
        self.radius = D.from_numpy(np.array([self.st.radius], np.float32))
        self.vox, self.vox_off, self.status = D((n, 3)), D(2, np.uint32), D(1, np.uint32)
        self.vx, self.vy, self.vz = D(n), D(n), D(n)
        self.sor, self.sor_off = D((n, 3)), D(2, np.uint32)
        self.sx, self.sy, self.sz = D(n), D(n), D(n)
        self.nx, self.ny, self.nz = D(n), D(n), D(n)
        self.means = D(n) if keep_means else None
        self.model = np.zeros(4, np.float32)
        self.inl = np.zeros(1, np.int32)
        self.its = np.zeros(1, np.int32)

    def step(self):
        st, s, n = self.st, self.s, self.n
        call("rvp_batch_voxel", *self.frame.ptrs(), self.off.ptr, 1, n, self.leaf.ptr, self.vox.ptr,
             self.vox_off.ptr, self.status.ptr, s)
        # what the stage kept, counted on the device (the second entry of its output offsets)
        call("rvp_aos_to_soa_bounded", self.vox.ptr, n, self.vox_off.ptr + 4, self.vx.ptr, self.vy.ptr, self.vz.ptr, s)
        call("rvp_batch_sor_exact", self.vx.ptr, self.vy.ptr, self.vz.ptr, self.vox_off.ptr, 1, n, st.sor_k,
             st.sor_alpha, self.sor.ptr, self.sor_off.ptr, None, self.means.ptr if self.means else None, s)
        call("rvp_aos_to_soa_bounded", self.sor.ptr, n, self.sor_off.ptr + 4, self.sx.ptr, self.sy.ptr, self.sz.ptr, s)
        kept = (self.sx.ptr, self.sy.ptr, self.sz.ptr, self.sor_off.ptr, 1, n)
        call("rvp_batch_normals_radius", *kept, self.radius.ptr, None, st.max_nn, *st.vp, self.nx.ptr,
             self.ny.ptr, self.nz.ptr, s)
        P = C.POINTER
        call("rvp_batch_ransac_plane", *kept, st.thr, st.iters, st.collinear, st.probability,
             self.model.ctypes.data_as(P(C.c_float)), self.inl.ctypes.data_as(P(C.c_int)),
             self.its.ctypes.data_as(P(C.c_int)), s)

    def info(self):
        return {"n_in": self.n, "n_voxel": int(self.vox_off.numpy()[1]), "n_sor": int(self.sor_off.numpy()[1]),
                "inliers": int(self.inl[0]), "ransac_iters": int(self.its[0]),
                "model": [float(v) for v in self.model]}

    def results(self):
        """Host copies of every stage's output: voxel centroids, SOR means (when kept) and kept points, normals
        of the kept points, the plane."""
        nv = int(self.vox_off.numpy()[1])
        ns = int(self.sor_off.numpy()[1])
        out = {"voxels": self.vox.numpy()[:nv], "kept": self.sor.numpy()[:ns],
               "normals": np.stack([self.nx.numpy()[:ns], self.ny.numpy()[:ns], self.nz.numpy()[:ns]], 1),
               "model": self.model.copy(), "inliers": int(self.inl[0]), "iterations": int(self.its[0])}
        if self.means is not None:
            out["means"] = self.means.numpy()[:nv]
        return out
