"""Stage chains on the device.

DeviceFramePipeline: the LiDAR-frame chain, voxel grid -> exact-kNN SOR -> radius normals -> RANSAC plane.

The chain follows the reference's pipeline driver (src/tools/pipeline_export.cpp:356-480: voxel_grid_downsamp,
sor, normal_estimation with a radius search, ransac_plane), with a cap on the radius neighbours.  The frame runs
as a batch of one cloud: every stage reads its point count from the previous stage's device offsets, so nothing
waits on the host until RANSAC's adaptive stop is replayed at the end.

DeviceBatchPipeline: many clouds back to back in HBM, one launch per stage: exact-kNN SOR -> capped radius lists
of the kept points -> k-NN normals of the kept points.
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

    keep_means: also keep SOR's per-point mean k-NN distances (voxel order) for inspection.
    graph: replay voxel -> SOR -> normals as one captured HIP graph (the kernels only; RANSAC's adaptive loop
    stays on the host); needs a created stream, which the pipeline makes when none is given.  The first step
    runs eagerly and captures the graph; later steps replay it.  n_max: device capacity in points (default the
    frame's); load() puts a new frame of at most n_max points into the same buffers."""

    def __init__(self, frame_xyz, stages: FrameStages | None = None, stream=None, keep_means=False, graph=False,
                 n_max=None):
        from .device import Stream
        self.st = stages or FrameStages()
        self._own_stream = None
        if graph and stream is None:
            self._own_stream = Stream()
            stream = self._own_stream.handle
        self.s = stream
        self.use_graph = graph
        self.graph = None
        self.n_max = n = int(n_max or len(frame_xyz))
        D = DeviceArray
        self.frame = DeviceCloud(np.zeros((n, 3), np.float32), stream)
        self.off = D(2, np.uint32)
        self.leaf = D.from_numpy(np.array([self.st.leaf], np.float32), stream)
        self.radius = D.from_numpy(np.array([self.st.radius], np.float32), stream)
        self.vox, self.vox_off, self.status = D((n, 3)), D(2, np.uint32), D(1, np.uint32)
        self.vx, self.vy, self.vz = D(n), D(n), D(n)
        self.sor, self.sor_off = D((n, 3)), D(2, np.uint32)
        self.sx, self.sy, self.sz = D(n), D(n), D(n)
        self.nx, self.ny, self.nz = D(n), D(n), D(n)
        self.means = D(n) if keep_means else None
        self.model = np.zeros(4, np.float32)
        self.inl = np.zeros(1, np.int32)
        self.its = np.zeros(1, np.int32)
        self.load(frame_xyz)

    def load(self, frame_xyz):
        """A new frame (at most n_max points) into the pipeline's input buffers; a captured graph reads it."""
        xyz = np.ascontiguousarray(frame_xyz, np.float32).reshape(-1, 3)
        if len(xyz) > self.n_max:
            raise ValueError(f"frame of {len(xyz)} points exceeds the pipeline's {self.n_max}")
        self.n = len(xyz)
        for dst, i in ((self.frame.x, 0), (self.frame.y, 1), (self.frame.z, 2)):
            col = np.ascontiguousarray(xyz[:, i])
            if len(col):
                call("rvp_memcpy_h2d", dst.ptr, col.ctypes.data_as(C.c_void_p), col.nbytes, self.s)
        off = np.array([0, self.n], np.uint32)
        call("rvp_memcpy_h2d", self.off.ptr, off.ctypes.data_as(C.c_void_p), off.nbytes, self.s)

    def _device_stages(self):
        st, s, n = self.st, self.s, self.n_max
        call("rvp_batch_voxel", *self.frame.ptrs(), self.off.ptr, 1, n, self.leaf.ptr, self.vox.ptr,
             self.vox_off.ptr, self.status.ptr, s)
        # what the stage kept, counted on the device (the second entry of its output offsets)
        call("rvp_aos_to_soa_bounded", self.vox.ptr, n, self.vox_off.ptr + 4, self.vx.ptr, self.vy.ptr, self.vz.ptr, s)
        call("rvp_batch_sor_exact", self.vx.ptr, self.vy.ptr, self.vz.ptr, self.vox_off.ptr, 1, n, st.sor_k,
             st.sor_alpha, self.sor.ptr, self.sor_off.ptr, None, self.means.ptr if self.means else None, s)
        call("rvp_aos_to_soa_bounded", self.sor.ptr, n, self.sor_off.ptr + 4, self.sx.ptr, self.sy.ptr, self.sz.ptr, s)
        call("rvp_batch_normals_radius", self.sx.ptr, self.sy.ptr, self.sz.ptr, self.sor_off.ptr, 1, n,
             self.radius.ptr, None, st.max_nn, *st.vp, self.nx.ptr, self.ny.ptr, self.nz.ptr, s)

    def _ransac(self, stream=None):
        st, s = self.st, (self.s if stream is None else stream)
        P = C.POINTER
        call("rvp_batch_ransac_plane", self.sx.ptr, self.sy.ptr, self.sz.ptr, self.sor_off.ptr, 1, self.n_max,
             st.thr, st.iters, st.collinear, st.probability, self.model.ctypes.data_as(P(C.c_float)),
             self.inl.ctypes.data_as(P(C.c_int)), self.its.ctypes.data_as(P(C.c_int)), s)

    def _capture(self):
        g = C.c_void_p()
        call("rvp_graph_capture_begin", self.s)
        try:
            self._device_stages()
        finally:
            call("rvp_graph_capture_end", self.s, C.byref(g))
        self.graph = g.value

    def prepare(self):
        """Graph mode: the eager first run (first-use allocations happen outside the capture) and the capture;
        nothing otherwise, or once done."""
        if self.use_graph and self.graph is None:
            self._device_stages()
            self._ransac()
            self._capture()
            call("rvp_stream_sync", self.s)

    def launch(self):
        """Enqueue voxel -> SOR -> normals on the pipeline's stream (the captured graph when there is one)."""
        self.prepare()
        if self.graph is not None:
            call("rvp_graph_launch", self.graph, self.s)
        else:
            self._device_stages()

    def step(self):
        self.launch()
        self._ransac()

    def sync(self):
        call("rvp_stream_sync", self.s)

    def info(self):
        return {"n_in": self.n, "n_voxel": int(self.vox_off.numpy(stream=self.s)[1]),
                "n_sor": int(self.sor_off.numpy(stream=self.s)[1]), "inliers": int(self.inl[0]),
                "ransac_iters": int(self.its[0]), "model": [float(v) for v in self.model]}

    def results(self):
        """Host copies of every stage's output: voxel centroids, SOR means (when kept) and kept points, normals
        of the kept points, the plane."""
        s = self.s
        nv = int(self.vox_off.numpy(stream=s)[1])
        ns = int(self.sor_off.numpy(stream=s)[1])
        out = {"voxels": self.vox.numpy(stream=s)[:nv], "kept": self.sor.numpy(stream=s)[:ns],
               "normals": np.stack([a.numpy(stream=s)[:ns] for a in (self.nx, self.ny, self.nz)], 1),
               "model": self.model.copy(), "inliers": int(self.inl[0]), "iterations": int(self.its[0])}
        if self.means is not None:
            out["means"] = self.means.numpy(stream=s)[:nv]
        return out

    def close(self):
        if self.graph is not None:
            from ._lib import lib
            lib().rvp_graph_destroy(self.graph)
            self.graph = None
        if self._own_stream is not None:
            self._own_stream.close()
            self._own_stream = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PipelinedFrames:
    """Frames through two DeviceFramePipelines in turn (double buffering), so that the device never waits for the
    host: each step enqueues its frame's voxel -> SOR -> normals (a captured graph) on the main stream, records an
    event behind it, and only then finishes the previous frame, whose RANSAC runs on a second stream after that
    frame's event while the host replays the adaptive stop (ransac_plane.cpp:95-105).  The frame of step i
    therefore overlaps frame i - 1's RANSAC and host replay.  A frame's buffers are reused two steps later, after
    its RANSAC has returned.  drain() finishes the last frame; result(i % 2) / info are those of the pipelines.
    Every frame's results are those of DeviceFramePipeline.step() on the same input."""

    def __init__(self, frame_xyz, stages: FrameStages | None = None, graph=True, n_max=None, keep_means=False):
        from .device import Event, Stream
        self.stream, self.rstream = Stream(), Stream()
        self.pipes = [DeviceFramePipeline(frame_xyz, stages, stream=self.stream.handle, graph=graph, n_max=n_max,
                                          keep_means=keep_means) for _ in range(2)]
        self.events = [Event(), Event()]
        for p in self.pipes:
            p.prepare()
        self.i = 0
        self.pending = None  # pipeline index whose frame awaits RANSAC
        self.finished = []   # (frame number, pipeline index) in finishing order, for callers that track frames

    def load(self, frame_xyz):
        """The next step's frame (into the pipeline that step uses; its previous frame is finished already)."""
        self.pipes[self.i % 2].load(frame_xyz)

    def step(self):
        k = self.i % 2
        self.pipes[k].launch()
        self.events[k].record(self.stream.handle)
        self._finish()
        self.pending = (self.i, k)
        self.i += 1

    def _finish(self):
        if self.pending is None:
            return
        n, k = self.pending
        call("rvp_stream_wait_event", self.rstream.handle, self.events[k].handle)
        self.pipes[k]._ransac(self.rstream.handle)
        self.pending = None
        self.finished.append((n, k))

    def drain(self):
        self._finish()
        self.stream.sync()
        self.rstream.sync()

    def last(self):
        """The pipeline holding the most recently finished frame."""
        return self.pipes[self.finished[-1][1]] if self.finished else None

    def close(self):
        self.drain()
        for p in self.pipes:
            p.close()
        for e in self.events:
            e.close()
        self.stream.close()
        self.rstream.close()


class ConcurrentFrames:
    """Frames on n DeviceFramePipelines at once, each with its own buffers, stream and captured graph and driven
    by its own host thread: n frames in flight, whose kernels the device interleaves.  A frame alone leaves much
    of the GPU idle (look-back chains, one-wave sums, launch gaps, RANSAC's host round trip); the other frames'
    kernels fill it.  Every frame's results are those of DeviceFramePipeline.step() on the same input (the
    pipelines share no buffers; the library's per-stream state keeps their launches apart).

    run(frames_per_pipe) -> (wall seconds, per-frame latencies in ms: start of a frame's step to its RANSAC result
    on the host).  load(i, xyz) puts a frame into pipeline i."""

    def __init__(self, frame_xyz, n=2, stages: FrameStages | None = None, n_max=None, keep_means=False):
        self.pipes = [DeviceFramePipeline(frame_xyz, stages, graph=True, n_max=n_max, keep_means=keep_means)
                      for _ in range(n)]
        for p in self.pipes:
            p.prepare()

    def load(self, i, frame_xyz):
        self.pipes[i].load(frame_xyz)

    def run(self, frames_per_pipe, on_frame=None):
        """Each pipeline runs frames_per_pipe steps in its own thread (the ctypes calls release the GIL);
        on_frame(i, j), when given, is called by pipeline i's thread after its frame j."""
        import threading
        import time
        lat = [[] for _ in self.pipes]
        errors = []

        def work(i):
            p = self.pipes[i]
            try:
                for j in range(frames_per_pipe):
                    t0 = time.perf_counter()
                    p.step()
                    lat[i].append((time.perf_counter() - t0) * 1e3)
                    if on_frame is not None:
                        on_frame(i, j)
            except Exception as e:  # noqa: BLE001 -- re-raised in the caller's thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(i,)) for i in range(len(self.pipes))]
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for p in self.pipes:
            p.sync()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return wall, [x for l in lat for x in l]

    def close(self):
        for p in self.pipes:
            p.sync()
            p.close()


class DeviceBatchPipeline:
    """Clouds back to back in HBM, one launch per stage and no host sync between stages: exact-kNN SOR (k_sor)
    -> radius lists (radius per cloud, capped at cap) of the kept points -> k-NN normals (k_nrm) of the kept
    points, taken from the radius lists where a list holds the k_nrm + 1 nearest (the same normals as a search).  The defaults are the 256-cloud workload's: k=32, radius 0.02 * the cloud's scale, cap 64, k=24."""

    K_SOR, K_NRM, CAP = 32, 24, 64

    def __init__(self, clouds, scales, stream=None):
        self.s = stream
        self.sizes = np.array([len(c) for c in clouds], np.int64)
        self.ncl = len(clouds)
        self.off_host = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.uint32)
        self.n = int(self.off_host[-1])
        n, D = self.n, DeviceArray
        self.cloud = DeviceCloud(np.concatenate(clouds))
        self.off = D.from_numpy(self.off_host)
        self.radius_host = (0.02 * np.asarray(scales)).astype(np.float32)
        self.radius = D.from_numpy(self.radius_host)
        self.means = D(n)
        self.sor, self.sor_off = D((n, 3)), D(self.ncl + 1, np.uint32)
        self.sx, self.sy, self.sz = D(n), D(n), D(n)
        self.idx, self.d2, self.cnt = D((n, self.CAP), np.int32), D((n, self.CAP)), D(n, np.int32)
        self.nx, self.ny, self.nz = D(n), D(n), D(n)

    def step(self):
        s, n = self.s, self.n
        call("rvp_batch_sor_exact", *self.cloud.ptrs(), self.off.ptr, self.ncl, n, self.K_SOR, 1.0, self.sor.ptr,
             self.sor_off.ptr, None, self.means.ptr, s)
        # the kept total is the last entry of the batch's output offsets, read on the device
        call("rvp_aos_to_soa_bounded", self.sor.ptr, n, self.sor_off.ptr + 4 * self.ncl, self.sx.ptr, self.sy.ptr,
             self.sz.ptr, s)
        kept = (self.sx.ptr, self.sy.ptr, self.sz.ptr, self.sor_off.ptr, self.ncl, n)
        call("rvp_batch_radius_search", *kept, self.radius.ptr, None, self.CAP, self.idx.ptr, self.d2.ptr,
             self.cnt.ptr, s)
        # the k-NN normals read each query's K_NRM + 1 nearest from its radius list where that holds enough entries
        call("rvp_batch_normals_knn_lists", *kept, self.K_NRM, self.idx.ptr, self.d2.ptr, self.cnt.ptr, self.CAP,
             0.0, 0.0, 0.0, self.nx.ptr, self.ny.ptr, self.nz.ptr, s)

    def kept_total(self):
        return int(self.sor_off.numpy(stream=self.s)[-1])
