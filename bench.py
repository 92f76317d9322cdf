#!/usr/bin/env python3
"""Single LiDAR frame through the neighbour-search pipeline on one MI355X.

Workload (one "step"): a 64-ring x 18,750-azimuth LiDAR frame (1.2M rays, ~1.04M returns, resident in HBM)
  voxel grid 0.1 m -> exact-kNN SOR k=50 alpha=1 -> normals, radius 0.5 m capped at 64 neighbours
  -> RANSAC plane, 1000 hypotheses, threshold 0.05 m
Metric: input points per second of the whole pipeline (Mpts/s), with p50/p99 ms per frame, for the frame run
as a batch of one (stages chained on device-resident counts; the value) and through the single-cloud entry
points with a host round trip per stage (plain_path); both must give the same voxel/SOR counts and plane.
Also reported: the dominant kernel's HBM-roofline fraction (HIP events around its launches on its own
stream), the 256-cloud batched vs per-cloud leg, and the reference's scalar path timed on this host
(oracle/_ref, or the oracle port).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 VALU instruction per SIMD every 4 cycles
VALU_PEAK_WAVE_INSTR_S = 256 * 4 * 2.4e9 / 4
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lidar_frame_hbm_traffic.txt")

# Algorithmic HBM bytes per unit of work for each stage (compulsory traffic: inputs read once, outputs
# written once).  unit = input point of the stage unless noted.
STAGE_BYTES = {
    "voxel": (20, 16),          # 12 read + 8 key per point, + 16 per voxel (centroid + count)
    "sor": (41, 0),             # 12 query + 12 candidate + 4 mean + 1 mask + 12 * keep (upper bound)
    "hash": (28, 16),           # 12 read + 12 sorted write + 4 index, + 16 per occupied cell
    "normals": (36, 0),         # 12 query + 12 candidate + 12 normal
    "ransac": (12, 0),          # 12 per point per launch evaluating all its hypotheses
}
KERNEL_STAGE = {
    "voxel_keys_k": "voxel", "voxel_centroid_k": "voxel", "voxel_minmax_k": "voxel", "seg_vox_keys_k": "voxel",
    "seg_vox_minmax_k": "voxel",
    "sor_knn_means_k": "sor", "sor_stats_k": "sor", "sor_keep_k": "sor", "seg_sor_means_k": "sor",
    "keys_k": "hash", "gather_k": "hash", "cell_start_k": "hash", "seg_keys_k": "hash", "seg_gather_k": "hash",
    "seg_cell_start_k": "hash", "seg_fill_gaps_k": "hash",
    "normals_topk_k": "normals", "seg_normals_k": "normals", "plane_count_k": "ransac", "seg_plane_count_k": "ransac",
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sor-sample", type=int, default=0,
                    help="voxels in the O(N^2) SOR CPU sample (0: all, nothing extrapolated)")
    ap.add_argument("--cpu-normals-sample", type=int, default=0, help="points in the CPU normals sample (0: all)")
    ap.add_argument("--batch-steps", type=int, default=3, help="timed passes of the 256-cloud batch leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched vs per-cloud leg")
    ap.add_argument("--no-indoor", action="store_true", help="skip the 20M-point indoor leg")
    ap.add_argument("--indoor-steps", type=int, default=3)
    ap.add_argument("--no-serve", action="store_true", help="skip the streaming (serving) leg")
    ap.add_argument("--serve-frames", type=int, default=6)
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="one device-to-device copy of this many MiB before warmup (HBM counter calibration)")
    ap.add_argument("--only", choices=["frame", "batched", "indoor", "serving"], default=None,
                    help="run one leg alone (profiling runs); the JSON line then carries that leg only")
    ap.add_argument("--tune", action="append", default=[], metavar="NAME=VALUE",
                    help="set a library tuning knob before running (repeatable; rvp_tune_set names)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the all-core CPU leg (0: all this "
                    "process may use)")
    return ap.parse_args()


class FramePipeline:
    """The plain GPU path: voxel -> SOR -> normals -> RANSAC through the single-cloud entry points, one frame
    per step; every stage hands its point count back to the host before the next one starts."""

    def __init__(self, frame_xyz, stream=None):
        import rvpoint_amd as R
        from rvpoint_amd._lib import call
        self.R, self.call, self.s = R, call, stream
        self.n = len(frame_xyz)
        n = self.n
        self.frame = R.DeviceCloud(frame_xyz)
        D = R.DeviceArray
        self.vox = D((n, 3)); self.vx, self.vy, self.vz = D(n), D(n), D(n)
        self.sor = D((n, 3)); self.sx, self.sy, self.sz = D(n), D(n), D(n)
        self.nx, self.ny, self.nz = D(n), D(n), D(n)
        h = C.c_void_p()
        call("rvp_hash_create", C.byref(h))
        self.hash = h.value
        self.leaf, self.k, self.alpha = 0.1, 50, 1.0
        self.radius, self.max_nn, self.vp = 0.5, 64, (0.0, 0.0, 1.7)
        self.thr, self.iters = 0.05, 1000

    def step(self):
        call, s = self.call, self.s
        nv, ns = C.c_size_t(), C.c_size_t()
        call("rvp_voxel", *self.frame.ptrs(), self.n, self.leaf, self.vox.ptr, C.byref(nv), s)
        call("rvp_aos_to_soa", self.vox.ptr, nv.value, self.vx.ptr, self.vy.ptr, self.vz.ptr, s)
        call("rvp_sor_exact", self.vx.ptr, self.vy.ptr, self.vz.ptr, nv.value, self.k, self.alpha, self.sor.ptr,
             C.byref(ns), None, None, s)
        call("rvp_aos_to_soa", self.sor.ptr, ns.value, self.sx.ptr, self.sy.ptr, self.sz.ptr, s)
        call("rvp_hash_build", self.hash, self.sx.ptr, self.sy.ptr, self.sz.ptr, ns.value, self.radius, s)
        call("rvp_normals_radius", self.hash, self.radius, self.max_nn, *self.vp, self.nx.ptr, self.ny.ptr,
             self.nz.ptr, s)
        model = (C.c_float * 4)()
        inl, its = C.c_int(), C.c_int()
        call("rvp_ransac_plane", self.sx.ptr, self.sy.ptr, self.sz.ptr, ns.value, self.thr, self.iters, 1e-6, 0.99,
             model, C.byref(inl), C.byref(its), s)
        return {"n_in": self.n, "n_voxel": nv.value, "n_sor": ns.value, "inliers": inl.value,
                "ransac_iters": its.value, "model": [float(v) for v in model]}


class PerCloudPipeline:
    """The same three stages one cloud per launch (the plain GPU path): SOR, host read of the kept count,
    hash build and radius lists, kNN normals, cloud after cloud on one stream."""

    def __init__(self, bp):
        from rvpoint_amd._lib import call
        self.bp, self.call = bp, call
        h = C.c_void_p()
        call("rvp_hash_create", C.byref(h))
        self.hash = h.value

    def run_cloud(self, c):
        bp, call, s = self.bp, self.call, self.bp.s
        o, m = int(bp.off_host[c]), int(bp.sizes[c])
        at = lambda arr, width=1, esz=4: arr.ptr + o * width * esz  # noqa: E731
        ns = C.c_size_t()
        call("rvp_sor_exact", *(p + 4 * o for p in bp.cloud.ptrs()), m, bp.K_SOR, 1.0, at(bp.sor, 3), C.byref(ns),
             None, None, s)
        k = ns.value
        call("rvp_aos_to_soa", at(bp.sor, 3), k, at(bp.sx), at(bp.sy), at(bp.sz), s)
        r = float(bp.radius_host[c])
        call("rvp_hash_build", self.hash, at(bp.sx), at(bp.sy), at(bp.sz), k, r, s)
        call("rvp_radius_search", self.hash, None, None, None, k, r, bp.CAP, at(bp.idx, bp.CAP), at(bp.d2, bp.CAP),
             at(bp.cnt), None, s)
        call("rvp_normals_knn", at(bp.sx), at(bp.sy), at(bp.sz), k, bp.K_NRM, 0.0, 0.0, 0.0, at(bp.nx), at(bp.ny),
             at(bp.nz), s)


def batch_leg(R, args):
    """Batched (one launch per stage) vs per-cloud on the 256-cloud workload: Mpts/s of input points."""
    from rvpoint_amd import scenes
    from rvpoint_amd.pipeline import DeviceBatchPipeline as BatchPipeline
    print("bench: generating 256-cloud batch", file=sys.stderr, flush=True)
    clouds, scales = scenes.batch_clouds(with_scales=True)
    bp = BatchPipeline(clouds, scales)
    del clouds
    for _ in range(max(args.warmup, 1)):
        bp.step()
    R.synchronize()
    bt = []
    for _ in range(args.batch_steps):
        t0 = time.perf_counter()
        bp.step()
        R.synchronize()
        bt.append((time.perf_counter() - t0) * 1e3)
    kept = bp.kept_total()
    # the batch keeps cloud c's kept points at kept offset c, the per-cloud leg at input offset c
    kept_off = bp.sor_off.numpy().astype(np.int64)
    batched_out = [a.numpy().view(np.uint32) for a in (bp.nx, bp.ny, bp.nz, bp.cnt)]
    print(f"bench: batched {np.median(bt):.2f} ms/batch; per-cloud leg", file=sys.stderr, flush=True)
    pc = PerCloudPipeline(bp)
    for c in range(bp.ncl):  # warm every cloud size once
        pc.run_cloud(c)
    R.synchronize()
    pt, per_cloud_ms = [], []
    for _ in range(args.batch_steps):
        t0 = time.perf_counter()
        for c in range(bp.ncl):
            t1 = time.perf_counter()
            pc.run_cloud(c)
            R.synchronize()
            per_cloud_ms.append((time.perf_counter() - t1) * 1e3)
        pt.append((time.perf_counter() - t0) * 1e3)
    # radius counts and normals of the per-cloud leg (same buffers) and of the batch agree bit for bit
    per_out = [a.numpy().view(np.uint32) for a in (bp.nx, bp.ny, bp.nz, bp.cnt)]
    same = all(np.array_equal(b[kept_off[c]:kept_off[c + 1]],
                              p[bp.off_host[c]:bp.off_host[c] + kept_off[c + 1] - kept_off[c]])
               for b, p in zip(batched_out, per_out) for c in range(bp.ncl))
    b_ms, p_ms = float(np.median(bt)), float(np.median(pt))
    return {
        "workload": "256 plane+sphere clouds, log-uniform 2e3..5e5 points, scale U[1,20]; SOR k=32 -> "
                    "radius r=0.02*scale cap 64 -> normals k=24",
        "points": bp.n, "kept_after_sor": kept, "clouds": bp.ncl, "steps": args.batch_steps,
        "batched_ms": round(b_ms, 3), "batched_p99_ms": round(float(np.percentile(bt, 99)), 3),
        "batched_mpts_s": round(bp.n / (b_ms * 1e-3) / 1e6, 3),
        "per_cloud_ms": round(p_ms, 3), "per_cloud_mpts_s": round(bp.n / (p_ms * 1e-3) / 1e6, 3),
        "per_cloud_p99_ms_per_cloud": round(float(np.percentile(per_cloud_ms, 99)), 3),
        "speedup": round(p_ms / b_ms, 3), "same_results": same,
    }


def time_steps(R, fn, steps):
    times = []
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        times.append((time.perf_counter() - t0) * 1e3)
    R.synchronize()
    return (time.perf_counter() - t_all) * 1e3 / steps, times


def indoor_leg(R, args):
    """Dense indoor scan, 20M points: hash build (cell 0.02), radius lists r=0.02 capped at 64 with indices
    and squared distances for every point, kNN normals k=30.  The radius kernel's HBM roofline uses its
    algorithmic bytes: 12 (query) + 12 (candidates once) + 4 (count) + 8 per returned neighbour."""
    from rvpoint_amd import _lib, scenes
    from rvpoint_amd._lib import call
    print("bench: generating 20M-point indoor scan", file=sys.stderr, flush=True)
    pts = scenes.indoor_scan()
    n = len(pts)
    D = R.DeviceArray
    dev = R.DeviceCloud(pts)
    del pts
    h = C.c_void_p()
    call("rvp_hash_create", C.byref(h))
    cap, r = 64, 0.02
    idx, d2, cnt = D((n, cap), np.int32), D((n, cap)), D(n, np.int32)
    nx, ny, nz = D(n), D(n), D(n)

    def build():
        call("rvp_hash_build", h.value, *dev.ptrs(), n, r, None)

    def radius():
        call("rvp_radius_search", h.value, None, None, None, n, r, cap, idx.ptr, d2.ptr, cnt.ptr, None, None)

    def normals():
        call("rvp_normals_knn", *dev.ptrs(), n, 30, 0.0, 0.0, 0.0, nx.ptr, ny.ptr, nz.ptr, None)

    # the same k=30 normals with each point's 31 nearest read from its radius list when it holds 31 entries
    off = D.from_numpy(np.array([0, n], np.uint32))
    lx, ly, lz = D(n), D(n), D(n)

    def normals_lists():
        call("rvp_batch_normals_knn_lists", *dev.ptrs(), off.ptr, 1, n, 30, idx.ptr, d2.ptr, cnt.ptr, cap, 0.0, 0.0,
             0.0, lx.ptr, ly.ptr, lz.ptr, None)

    out = {"workload": "indoor scan 20M points (10x8x3 m room, 60 boxes, 20 cylinders, sigma 3 mm): hash build "
                       "cell 0.02 -> radius lists r=0.02 cap 64 (idx + d2, every point) ; normals k=30 searched, and read "
                       "from the radius lists where a list holds the 31 nearest",
           "points": n}
    for name, fn, kern in (("hash_build", build, None), ("radius", radius, "radius_topk_k"),
                           ("normals_k30", normals, "normals_topk_k"),
                           ("normals_k30_lists", normals_lists, "seg_normals_k")):
        fn()
        R.synchronize()
        if kern:
            _lib.profile_enable(True, kern)
        ms, _ = time_steps(R, fn, args.indoor_steps)
        if kern:
            _lib.profile_enable(False)
            prof = _lib.profile_collect()
            launches, total = prof.get(kern, (0, 0.0))
            out[name + "_kernel_ms"] = round(total / max(launches, 1), 3)
        out[name + "_ms"] = round(ms, 3)
        out[name + "_mpts_s"] = round(n / (ms * 1e-3) / 1e6, 2)
        if name == "hash_build":
            build()  # the hash the queries use
    m = cnt.numpy().astype(np.int64)
    mean_nn = float(m.mean())
    out["normals_k30_lists_served"] = round(float((m >= 31).mean()), 4)
    out["normals_k30_lists_same"] = all(np.array_equal(a.numpy().view(np.uint32), b.numpy().view(np.uint32))
                                        for a, b in ((nx, lx), (ny, ly), (nz, lz)))
    bytes_per_launch = n * (12 + 12 + 4) + 8 * int(m.sum())
    kms = out["radius_kernel_ms"]
    achieved = bytes_per_launch / (kms * 1e-3) / 1e9
    out["radius_mean_neighbours"] = round(mean_nn, 2)
    out["radius_roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(achieved / HBM_PEAK_GBS, 5), "bytes_per_launch": bytes_per_launch}
    call("rvp_hash_destroy", h.value)
    return out


def serving_leg(R, args):
    """Streaming LiDAR-like frames (rings in {16, 32, 64, 128}) through the frame server (continuous batching,
    admission by an HBM budget): capacity under a burst, then Poisson arrivals at fractions of it with
    frames/s, Mpts/s and p50/p99 latency from submission to results."""
    from rvpoint_amd import scenes
    from rvpoint_amd.serve import FrameParams, FrameServer, poisson_run
    plan = scenes.lidar_stream(frames=args.serve_frames)
    print(f"bench: generating {len(plan)} streaming frames", file=sys.stderr, flush=True)
    frames = [scenes.lidar_frame(rings=r, seed=sd) for r, sd in plan]
    sizes = [len(f) for f in frames]
    srv = FrameServer(FrameParams.lidar(), slots=args.serve_slots, slot_points=4_000_000, hbm_budget=64 << 30)
    try:
        for f in frames:  # warm every size once
            srv.wait(srv.submit(f))
        burst = frames * 3
        # one untimed burst: the pinned staging and result buffers a burst holds at once are then pooled
        for t in [srv.submit(f) for f in burst]:
            srv.wait(t)
        t0 = time.perf_counter()
        res = [srv.wait(t) for t in [srv.submit(f) for f in burst]]
        dt = time.perf_counter() - t0
        cap_fps = len(burst) / dt
        out = {"workload": "LiDAR-like frames, rings in {16,32,64,128} x 18750 azimuths, voxel 0.1 -> SOR k=50 -> "
                           "normals r=0.5 cap 64 -> RANSAC 1000; continuous batching over "
                           f"{srv.slots} slots of 4M points, HBM budget 64 GiB",
               "frame_points": sizes, "slots": srv.slots,
               "burst": {"frames": len(burst), "frames_s": round(cap_fps, 2),
                         "mpts_s": round(sum(r["n_in"] for r in res) / dt / 1e6, 2),
                         "mean_batch_frames": round(float(np.mean([r["batch_frames"] for r in res])), 2)},
               "offered": []}
        for frac in (0.3, 0.6, 0.9):
            rate = frac * cap_fps
            done = poisson_run(srv, frames, rate, args.serve_seconds)
            lat = np.array([(r["t_done"] - r["t_submit"]) * 1e3 for r in done])
            span = max(r["t_done"] for r in done) - min(r["t_submit"] for r in done)
            out["offered"].append({"rate_frames_s": round(rate, 2), "frames": len(done),
                                   "frames_s": round(len(done) / span, 2),
                                   "mpts_s": round(sum(r["n_in"] for r in done) / span / 1e6, 2),
                                   "p50_ms": round(float(np.percentile(lat, 50)), 3),
                                   "p99_ms": round(float(np.percentile(lat, 99)), 3),
                                   "mean_batch_frames": round(float(np.mean([r["batch_frames"] for r in done])), 2)})
            print(f"bench: offered {rate:.1f} frames/s -> {out['offered'][-1]}", file=sys.stderr, flush=True)
        return out
    finally:
        srv.close()


def cpu_baseline(frame, info, args):
    """The reference's scalar path (oracle/_ref when present, else the oracle port) on one host core: by
    default the whole frame, stage by stage, each stage on the previous stage's output (voxel 0.1 -> sor_sc
    k=50 -> SpatialHash radius-0.5 normals -> ransac 1000), nothing extrapolated.  --cpu-sor-sample /
    --cpu-normals-sample bound the two expensive stages to a subset, scaled by N^2 and N."""
    import oracle as O
    kind = "reference" if O.ref_available() else "port"
    pts = np.ascontiguousarray(frame, np.float32)
    n = len(pts)
    t = {}
    if kind == "reference":
        R = O.ref()
        out = np.zeros_like(pts)
        t0 = time.perf_counter()
        nv = R.ref_voxel_sc(pts.reshape(-1), n, out.reshape(-1), 0.1)
        t["voxel"] = time.perf_counter() - t0
        vox = out[:nv].copy()
    else:
        t0 = time.perf_counter()
        vox = O.voxel_sc(pts, 0.1)
        t["voxel"] = time.perf_counter() - t0
        nv = len(vox)
    m = nv if args.cpu_sor_sample <= 0 else min(args.cpu_sor_sample, nv)
    sub = vox if m == nv else np.ascontiguousarray(vox[np.random.default_rng(0).choice(nv, m, replace=False)])
    t0 = time.perf_counter()
    if kind == "reference":
        o = np.zeros_like(sub)
        ms = O.ref().ref_sor_sc(sub.reshape(-1), m, o.reshape(-1), 50, 1.0)
        kept = o[:ms].copy()
    else:
        kept, _ = O.sor_sc(sub, 50, 1.0)
    t["sor_sample"] = time.perf_counter() - t0
    # the all-pairs distance pass is N^2; the scaled figure is an estimate (the default measures it all)
    t["sor"] = t["sor_sample"] * (nv / m) ** 2
    # later stages run on the SOR output when SOR ran on the whole cloud, else on the first n_sor voxels
    stage_in = kept if m == nv else vox[:info["n_sor"]]
    ns = len(stage_in)
    m2 = ns if args.cpu_normals_sample <= 0 else min(args.cpu_normals_sample, ns)
    sx, sy, sz = (np.ascontiguousarray(stage_in[:m2, i]) for i in range(3))
    t0 = time.perf_counter()
    if kind == "reference":
        nx, ny, nz = (np.zeros(m2, np.float32) for _ in range(3))
        O.ref().ref_normal_hash(sx, sy, sz, m2, 0.5, nx, ny, nz, 10, 0.5, 0.0, 0.0, 1.7, 4)
    else:
        O.normals_radius_f64(sx, sy, sz, 0.5, 0.5, 64, (0.0, 0.0, 1.7))
    t["normals_sample"] = time.perf_counter() - t0
    t["normals"] = t["normals_sample"] * ns / m2
    x, y, z = (np.ascontiguousarray(stage_in[:, i]) for i in range(3))
    t0 = time.perf_counter()
    if kind == "reference":
        model = np.zeros(4, np.float32)
        O.ref().ref_ransac_rvv(x, y, z, ns, 0.05, 1000, model, 1e-6, 0.99)
    else:
        O.ransac_plane(x, y, z, 0.05, 1000)
    t["ransac"] = time.perf_counter() - t0
    total = t["voxel"] + t["sor"] + t["normals"] + t["ransac"]
    sampled = t["voxel"] + t["sor_sample"] + t["normals_sample"] + t["ransac"]
    if m == nv and m2 == ns:
        sample = (f"whole frame through the reference scalar path, one thread: voxel 0.1 ({n} -> {nv} pts), sor_sc "
                  f"k=50 ({nv} -> {ns}), SpatialHash radius-0.5 normals (uncapped, {ns} pts), ransac 1000 hyps; "
                  f"{total:.1f} s measured, nothing extrapolated")
    else:
        sample = (f"voxel 0.1 full frame ({n} pts); sor_sc k=50 on {m} random voxels scaled by (n/m)^2 to {nv}; "
                  f"SpatialHash radius-0.5 normals on {m2} pts scaled linearly to {ns}; ransac 1000 hyps; "
                  f"{sampled:.1f} s measured, {total:.1f} s extrapolated frame time")
    return {"value": round(n / total / 1e6, 6), "unit": "Mpts/s", "cores": 1, "kind": kind, "sample": sample,
            "stage_seconds": {k: round(v, 4) for k, v in t.items()}}


def host_cpu():
    """CPUs this process may use and the host's CPU model name."""
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return ncpu, model


def cpu_all_core(frame, gpu_info, args):
    """The same frame through the oracle's restatement of the scalar path on every host core: the per-point work
    (SOR's k nearest and mean, the hash normals) in OpenMP threads over points, the global SOR sums, voxel sums
    and the RANSAC loop sequential as in the reference; its outputs equal the reference's bit for bit (the
    threaded functions are pinned against oracle/_ref on the CPU)."""
    import oracle as O
    ncpu, model = host_cpu()
    th = args.cpu_threads or ncpu
    pts = np.ascontiguousarray(frame, np.float32)
    n = len(pts)
    t = {}
    t0 = time.perf_counter()
    vox = O.voxel_sc(pts, 0.1)
    t["voxel"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    kept, _ = O.sor_sc_mt(vox, 50, 1.0, nthreads=th)
    t["sor"] = time.perf_counter() - t0
    x, y, z = (np.ascontiguousarray(kept[:, i]) for i in range(3))
    t0 = time.perf_counter()
    O.normals_hash_e1(x, y, z, 0.5, 0.5, (0.0, 0.0, 1.7), 4, nthreads=th)
    t["normals"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.ransac_plane(x, y, z, 0.05, 1000)
    t["ransac"] = time.perf_counter() - t0
    total = sum(t.values())
    return {"value": round(n / total / 1e6, 6), "unit": "Mpts/s", "cores": th, "kind": "port",
            "host_cpus": ncpu, "cpu_model": model,
            "sample": f"whole frame, oracle port on {th} threads (OpenMP over points; global sums, voxel and RANSAC "
                      f"sequential): voxel 0.1 ({n} -> {len(vox)}), sor_sc k=50 ({len(vox)} -> {len(kept)}), "
                      f"hash radius-0.5 normals with the reference's eigen step, ransac 1000; {total:.2f} s",
            "same_counts_as_gpu": bool(len(vox) == gpu_info["n_voxel"] and len(kept) == gpu_info["n_sor"]),
            "stage_seconds": {k: round(v, 4) for k, v in t.items()}}


def frame_precision(pipe):
    """The frame's normals (radius 0.5, cap 64, viewpoint (0, 0, 1.7)) against the FP64 oracle's normals of the
    same capped (d2, idx) neighbourhoods: median / p99 / max angle in degrees and the FP64-fallback count."""
    import oracle as O
    from rvpoint_amd import normals_fallback_count
    res = pipe.results()
    kept, got = res["kept"], res["normals"].astype(np.float64)
    x, y, z = (np.ascontiguousarray(kept[:, i]) for i in range(3))
    ref = O.normals_radius_f64(x, y, z, 0.5, 0.5, 64, (0.0, 0.0, 1.7))
    ok = np.linalg.norm(ref, axis=1) > 0
    c = np.abs((got[ok] * ref[ok]).sum(1)) / np.maximum(np.linalg.norm(got[ok], axis=1), 1e-300)
    ang = np.degrees(np.arccos(np.clip(c, -1.0, 1.0)))
    return {"normals_vs_fp64_deg": {"median": float(np.median(ang)), "p99": float(np.percentile(ang, 99)),
                                    "max": float(ang.max()), "points": int(ok.sum())},
            "fp64_fallbacks": int(normals_fallback_count()), "zero_normal_sets_equal":
            bool(((np.linalg.norm(got, axis=1) == 0) == ~ok).all())}


def copy_bandwidth(R, mib=1024, reps=5):
    """Measured device-to-device copy bandwidth (read + write bytes per second), HIP events on one stream."""
    from rvpoint_amd._lib import call
    nb = mib << 20
    a, b = R.DeviceArray(nb // 4), R.DeviceArray(nb // 4)
    ev = [C.c_void_p(), C.c_void_p()]
    for e in ev:
        call("rvp_event_create", C.byref(e))
    call("rvp_memcpy_d2d", b.ptr, a.ptr, nb, None)
    best = None
    for _ in range(reps):
        call("rvp_event_record", ev[0], None)
        call("rvp_memcpy_d2d", b.ptr, a.ptr, nb, None)
        call("rvp_event_record", ev[1], None)
        ms = C.c_float()
        call("rvp_event_elapsed", ev[0], ev[1], C.byref(ms))
        best = ms.value if best is None else min(best, ms.value)
    for e in ev:
        call("rvp_event_destroy", e)
    del a, b
    return round(2 * nb / (best * 1e-3) / 1e9, 1)


def measured_counters(kernel):
    """HBM bytes (FETCH_SIZE and WRITE_SIZE, scaled by a device copy of known size) and VALU instructions per
    dispatch of `kernel`, from the committed counter summary of this workload; (None, None) when absent."""
    if not kernel or not os.path.exists(TRAFFIC_FILE):
        return None, None
    for line in open(TRAFFIC_FILE):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        if f[0].split("<")[0] == kernel:
            return int(f[6]), int(f[7])
    return None, None


def roofline(prof, info, steps):
    """Dominant kernel by total time; achieved = algorithmic bytes per launch / mean launch duration."""
    if not prof:
        return None, {}
    name, (launches, total_ms) = max(prof.items(), key=lambda kv: kv[1][1])
    stage = KERNEL_STAGE.get(name)
    units = {"voxel": (info["n_in"], info["n_voxel"]), "sor": (info["n_voxel"], 0),
             "hash": (info["n_sor"], 0), "normals": (info["n_sor"], 0), "ransac": (info["n_sor"], 0)}
    if stage is None:
        return {"kernel": name, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}, prof
    per_pt, per_unit2 = STAGE_BYTES[stage]
    u1, u2 = units[stage]
    per_step_bytes = per_pt * u1 + per_unit2 * u2
    launches_per_step = launches / steps
    bytes_per_launch = per_step_bytes / launches_per_step
    avg_ms = total_ms / launches
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, valu = measured_counters(name)
    out = {"kernel": name, "bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
           "bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(avg_ms, 5),
           "stage": stage, "launches": launches}
    if traffic:
        out["traffic_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 3)
    if valu:
        # what actually bounds the query kernels: VALU issue (wave instructions per second against the peak)
        rate = valu / (avg_ms * 1e-3)
        out["valu"] = {"instr_per_launch": valu, "achieved": round(rate / 1e9, 2), "peak": VALU_PEAK_WAVE_INSTR_S / 1e9,
                       "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_WAVE_INSTR_S, 4)}
    return out, prof


def pipeline_roofline(info, per_kernel, ms):
    """The whole frame against HBM: algorithmic bytes of every stage (inputs read once, outputs written once) over
    the frame's wall time.  voxel 20 B/pt + 16 B/voxel; SOR 41 B/pt; the normals' grid 28 B/pt and normals
    36 B/pt; RANSAC 12 B/pt per counting launch."""
    rounds = per_kernel.get("seg_plane_count_k", (1, 0.0))[0]
    parts = {"voxel": 20 * info["n_in"] + 16 * info["n_voxel"], "sor": 41 * info["n_voxel"],
             "normals_grid": 28 * info["n_sor"], "normals": 36 * info["n_sor"],
             "ransac": int(12 * info["n_sor"] * rounds)}
    total = sum(parts.values())
    achieved = total / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "bytes_per_frame": int(total), "stage_bytes": parts, "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5)}


def frame_leg(R, args, frame):
    """The headline leg: the LiDAR frame through the plain path and the device frame pipeline."""
    from rvpoint_amd import _lib
    from rvpoint_amd.pipeline import DeviceFramePipeline
    # the plain path: single-cloud entry points, a host round trip per stage
    plain = FramePipeline(frame)
    for _ in range(args.warmup):
        plain.step()
    R.synchronize()
    plain_ms, plain_times = time_steps(R, plain.step, args.steps)
    info_plain = plain.step()
    R.synchronize()
    # the measured path: the frame as a batch of one, stages chained on device offsets
    pipe = DeviceFramePipeline(frame)
    for _ in range(args.warmup):
        pipe.step()
    R.synchronize()
    # which kernel dominates: two profiled steps outside the timed region
    _lib.profile_enable(True)
    for _ in range(2):
        pipe.step()
    R.synchronize()
    _lib.profile_enable(False)
    per_kernel = {k: (v[0] / 2, v[1] / 2) for k, v in _lib.profile_collect().items()}
    dominant = max(per_kernel.items(), key=lambda kv: kv[1][1])[0] if per_kernel else None
    print(f"bench: warmup done ({args.warmup} steps); dominant kernel {dominant}; timing {args.steps}",
          file=sys.stderr, flush=True)
    # timed region: HIP events only around the dominant kernel's launches, on the stream it runs on
    _lib.profile_enable(True, dominant)
    ms, times = time_steps(R, pipe.step, args.steps)
    _lib.profile_enable(False)
    prof = _lib.profile_collect()
    info = pipe.info()
    same = all(info[k] == info_plain[k] for k in ("n_voxel", "n_sor", "inliers", "model"))
    rl, prof = roofline(prof, info, args.steps)
    eager = {"ms_per_step": round(ms, 4), "p50_ms": round(float(np.percentile(times, 50)), 4),
             "p99_ms": round(float(np.percentile(times, 99)), 4), "mpts_s": round(info["n_in"] / (ms * 1e-3) / 1e6, 4)}
    # the same chain with voxel -> SOR -> normals replayed as one captured graph (RANSAC's loop stays eager)
    gpipe = DeviceFramePipeline(frame, graph=True)
    for _ in range(max(args.warmup, 1)):
        gpipe.step()
    R.synchronize()
    gms, gtimes = time_steps(R, gpipe.step, args.steps)
    ginfo = gpipe.info()
    gsame = all(ginfo[k] == info[k] for k in ("n_voxel", "n_sor", "inliers", "model"))
    graph = {"ms_per_step": round(gms, 4), "p50_ms": round(float(np.percentile(gtimes, 50)), 4),
             "p99_ms": round(float(np.percentile(gtimes, 99)), 4),
             "mpts_s": round(info["n_in"] / (gms * 1e-3) / 1e6, 4), "same_results": bool(gsame)}
    gpipe.close()
    # double-buffered frames: frame i + 1's graph is enqueued before frame i's RANSAC and its host replay, so the
    # device does not wait for the host between frames.  A frame's latency runs from the start of its step to
    # the return of the step that finishes it (the next one; drain() for the last frame).
    from rvpoint_amd.pipeline import PipelinedFrames
    pp = PipelinedFrames(frame)
    for _ in range(max(args.warmup, 1)):
        pp.step()
    pp.drain()
    starts, ends = [], []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        starts.append(time.perf_counter())
        pp.step()
        ends.append(time.perf_counter())
    pp.drain()
    t_end = time.perf_counter()
    pms = (t_end - t_all) * 1e3 / args.steps
    plat = [(ends[i + 1] - starts[i]) * 1e3 for i in range(args.steps - 1)] + [(t_end - starts[-1]) * 1e3]
    pinfo = pp.last().info()
    psame = all(pinfo[k] == info[k] for k in ("n_voxel", "n_sor", "inliers", "model"))
    pipelined = {"ms_per_step": round(pms, 4), "frame_latency_p50_ms": round(float(np.percentile(plat, 50)), 4),
                 "frame_latency_p99_ms": round(float(np.percentile(plat, 99)), 4),
                 "mpts_s": round(info["n_in"] / (pms * 1e-3) / 1e6, 4), "same_results": bool(psame)}
    pp.close()
    use_graph = gsame and gms < ms
    if use_graph:
        ms, times = gms, gtimes
    # the double-buffered form trades frame latency (about two frames) for a little throughput: reported, not
    # the headline, whose p50/p99 are per-frame latencies
    use_pipe = False
    # n frames in flight: n pipelines, each with its own stream, graph and host thread.  One frame leaves much of
    # the GPU idle (look-back chains, one-wave sums, launch gaps, RANSAC's host round trip), which the other
    # frames' kernels fill.  K steps = K frames, split over the pipelines; p50/p99 are per-frame latencies (the
    # start of a frame's step to its RANSAC result on the host) under that load.
    from rvpoint_amd.pipeline import ConcurrentFrames
    concurrent, conc_best = {}, None
    for nconc in (2, 3):
        cf = ConcurrentFrames(frame, n=nconc)
        cf.run(max(args.warmup, 1))
        fpp = -(-args.steps // nconc)
        wall, clat = cf.run(fpp)
        cms = wall * 1e3 / (fpp * nconc)
        csame = all(all(p.info()[k] == info[k] for k in ("n_voxel", "n_sor", "inliers", "model")) for p in cf.pipes)
        cf.close()
        concurrent[str(nconc)] = {"ms_per_frame": round(cms, 4), "frames": fpp * nconc,
                                  "frame_latency_p50_ms": round(float(np.percentile(clat, 50)), 4),
                                  "frame_latency_p99_ms": round(float(np.percentile(clat, 99)), 4),
                                  "mpts_s": round(info["n_in"] / (cms * 1e-3) / 1e6, 4), "same_results": bool(csame)}
        if nconc == 2 and csame:
            conc_best = (cms, clat)
    # the headline: two frames in flight when their results are the same and the throughput higher
    use_conc = conc_best is not None and conc_best[0] < ms
    if use_conc:
        ms, times = conc_best
    line = {
        "value": round(info["n_in"] / (ms * 1e-3) / 1e6, 4),
        "ms_per_step": round(ms, 4),
        "p50_ms": round(float(np.percentile(times, 50)), 4),
        "p99_ms": round(float(np.percentile(times, 99)), 4),
        "config": {"workload": "lidar_frame: 64 rings x 18750 azimuths (1.2M rays), voxel 0.1 -> SOR k=50 "
                               "-> normals r=0.5 cap 64 -> RANSAC 1000 thr 0.05",
                   "points": info["n_in"], "n_voxel": info["n_voxel"], "n_sor": info["n_sor"],
                   "inliers": info["inliers"], "parallelism": "single GPU",
                   "path": ("two frames in flight: two pipelines, each with its own stream, captured graph of "
                            "voxel -> SOR -> normals and host thread (RANSAC eager); " if use_conc else
                            "captured graph of voxel -> SOR -> normals, RANSAC eager; " if use_graph else "") +
                           "frame as a batch of one: stages chained on device offsets, one host wait (RANSAC)",
                   "p50_p99": ("frame latency under two frames in flight: start of the frame's step to its RANSAC "
                               "result on the host") if use_conc else "frame latency: one step per frame"},
        "device_eager": eager,
        "graph": graph,
        "pipelined": pipelined,
        "concurrent": concurrent,
        "plain_path": {"ms_per_step": round(plain_ms, 4), "p50_ms": round(float(np.percentile(plain_times, 50)), 4),
                       "p99_ms": round(float(np.percentile(plain_times, 99)), 4),
                       "mpts_s": round(info["n_in"] / (plain_ms * 1e-3) / 1e6, 4),
                       "same_results": bool(same)},
        "roofline": rl,
        "pipeline_roofline": pipeline_roofline(info, per_kernel, ms),
        "kernels_ms_per_step": {k: round(v[1], 4) for k, v in sorted(per_kernel.items(),
                                                                     key=lambda kv: -kv[1][1])[:12]},
    }
    return line, pipe, info


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU, and "
              "the multi-GPU split of this workload is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import rvpoint_amd as R
    from rvpoint_amd import _lib, scenes
    if not R.gpu_available():
        print("no GPU visible to librvp_hip.so", file=sys.stderr)
        sys.exit(1)
    from rvpoint_amd import autotune
    knobs_used = {}

    def tune_for(cls):
        """The autotuner's knobs for this workload class (tuning.json), then any --tune overrides."""
        autotune.apply(cls)
        for kv in args.tune:
            name, value = kv.split("=", 1)
            _lib.tune(name, float(value))
        knobs_used[cls] = {name: _lib.tune(name) for name in autotune.DEFAULTS}
    if args.copy_mib > 0:
        nb = args.copy_mib << 20
        a, b = R.DeviceArray(nb // 4), R.DeviceArray(nb // 4)
        _lib.call("rvp_memcpy_d2d", b.ptr, a.ptr, nb, None)
        R.synchronize()
        del a, b
    line = {"metric": "Mpts/s hash-grid radius/kNN+SOR+normals; batched vs per-cloud; p99 ms/frame", "value": None,
            "unit": "Mpts/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "tuning": knobs_used}
    legs = [args.only] if args.only else ["frame", "batched", "indoor", "serving"]
    frame = pipe = info = None
    if "frame" in legs:
        print("bench: generating LiDAR frame (64 rings x 18,750 azimuths)", file=sys.stderr, flush=True)
        frame = scenes.lidar_frame()
        tune_for("lidar_frame")
        part, pipe, info = frame_leg(R, args, frame)
        line.update(part)
        line["hbm_copy_gbs"] = copy_bandwidth(R)
    if "batched" in legs:
        tune_for("batch")
        line["batched"] = None if args.no_batch else batch_leg(R, args)
    if "indoor" in legs:
        tune_for("indoor")
        line["indoor"] = None if args.no_indoor else indoor_leg(R, args)
    if "serving" in legs:
        tune_for("lidar_frame")
        line["serving"] = None if args.no_serve else serving_leg(R, args)
    if "frame" in legs and not args.no_cpu_baseline:
        print("bench: frame precision against the FP64 oracle", file=sys.stderr, flush=True)
        line["precision"] = frame_precision(pipe)
        print("bench: CPU baseline (reference scalar path, 1 thread)", file=sys.stderr, flush=True)
        single = cpu_baseline(frame, info, args)
        ncpu, model = host_cpu()
        single["host_cpus"], single["cpu_model"] = ncpu, model
        print("bench: CPU baseline (all cores)", file=sys.stderr, flush=True)
        single["all_cores"] = cpu_all_core(frame, info, args)
        line["cpu_baseline"] = single
    elif "frame" in legs:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
