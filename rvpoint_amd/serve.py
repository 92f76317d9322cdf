"""Streaming frames through the device pipeline with continuous batching (the C server in server.hip).

    srv = FrameServer(FrameParams(), slots=3, slot_points=4_000_000, hbm_budget=64 << 30)
    t = srv.submit(frame_xyz)           # returns at once
    res = srv.wait(t)                   # dict: n_voxel, n_sor, points, normals, model, inliers, times
"""
from __future__ import annotations

import ctypes as C
import threading
import time

import numpy as np

from ._lib import call, lib


class FrameParams(C.Structure):
    _fields_ = [("leaf", C.c_float), ("sor_k", C.c_int), ("sor_alpha", C.c_float), ("normal_radius", C.c_float),
                ("normal_max_nn", C.c_int), ("vp", C.c_float * 3), ("ransac_thresh", C.c_float),
                ("ransac_iters", C.c_int), ("ransac_prob", C.c_float)]

    @classmethod
    def lidar(cls):
        """The LiDAR-frame pipeline: voxel 0.1 -> SOR k=50 -> normals r=0.5 cap 64 -> RANSAC 1000, thr 0.05."""
        return cls(0.1, 50, 1.0, 0.5, 64, (C.c_float * 3)(0.0, 0.0, 1.7), 0.05, 1000, 0.99)


class FrameResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("n_in", C.c_size_t), ("n_voxel", C.c_size_t), ("n_sor", C.c_size_t),
                ("points", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)), ("model", C.c_float * 4),
                ("inliers", C.c_int), ("iterations", C.c_int), ("voxel_status", C.c_uint32),
                ("batch_frames", C.c_int), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double)]


class FrameServer:
    def __init__(self, params: FrameParams | None = None, slots=2, slot_points=4_000_000, hbm_budget=64 << 30):
        self.params = params or FrameParams.lidar()
        h = C.c_void_p()
        call("rvp_server_create", C.byref(h), C.byref(self.params), int(slots), int(slot_points), int(hbm_budget))
        self.handle = h.value
        n = C.c_int()
        call("rvp_server_slots", self.handle, C.byref(n))
        self.slots = n.value

    def submit(self, xyz) -> int:
        a = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        t = C.c_uint64()
        call("rvp_server_submit", self.handle, a.ctypes.data_as(C.POINTER(C.c_float)), len(a), C.byref(t))
        return t.value

    def wait(self, ticket: int) -> dict:
        r = FrameResult()
        call("rvp_server_wait", self.handle, C.c_uint64(ticket), C.byref(r))
        try:
            m = r.n_sor
            pts = np.ctypeslib.as_array(r.points, (max(m, 1) * 3,))[:3 * m].reshape(m, 3).copy()
            nrm = np.ctypeslib.as_array(r.normals, (max(m, 1) * 3,))[:3 * m].reshape(m, 3).copy()
        finally:
            call("rvp_result_free", C.byref(r))
        return {"ticket": r.ticket, "n_in": r.n_in, "n_voxel": r.n_voxel, "n_sor": r.n_sor, "points": pts,
                "normals": nrm, "model": np.array(r.model, np.float32), "inliers": r.inliers,
                "iterations": r.iterations, "voxel_status": r.voxel_status, "batch_frames": r.batch_frames,
                "t_submit": r.t_submit, "t_start": r.t_start, "t_done": r.t_done}

    def close(self):
        if self.handle:
            lib().rvp_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_run(server: FrameServer, frames, rate_hz: float, duration_s: float, seed=4004):
    """Offer frames (cycled) with exponential inter-arrival times at rate_hz for duration_s; returns the
    per-frame results (without point arrays) in submission order."""
    rng = np.random.default_rng(seed)
    tickets = []
    done = []
    lock = threading.Lock()
    finished = threading.Event()

    def collect():
        i = 0
        while True:
            with lock:
                have = i < len(tickets)
                t = tickets[i] if have else None
            if not have:
                if finished.is_set():
                    with lock:
                        if i >= len(tickets):
                            return
                    continue
                time.sleep(0.0002)
                continue
            r = server.wait(t)
            r.pop("points")
            r.pop("normals")
            done.append(r)
            i += 1

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    nxt = t0
    j = 0
    while nxt - t0 < duration_s:
        now = time.perf_counter()
        if now < nxt:
            time.sleep(min(nxt - now, 0.002))
            continue
        tk = server.submit(frames[j % len(frames)])
        with lock:
            tickets.append(tk)
        j += 1
        nxt += rng.exponential(1.0 / rate_hz)
    finished.set()
    th.join()
    return done
