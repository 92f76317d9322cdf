"""Autotuner of the library's speed knobs, per workload class.

The knobs (rvp_tune_set) change only how fast the kernels run, never their results: the k-NN grid edge
(knn_cell_div: the estimated k-th neighbour distance over this), the dense cells a batched grid may give a point
(seg_cells_per_point), the query tile (queries_per_wave: consecutive queries one wave serves), how the batched
SOR means hand out work (work_steal), the first box of a k-NN search without a bound (knn_start_cells) and the
queries per wave of the radius normals (normals_per_wave).  One wave serves one
query at a time (64 lanes per query); that is fixed by the kernels' design.

tune() measures a workload class on the GPU by coordinate descent: each knob in turn over its candidate values,
the others held at their best so far, keeping the fastest setting (median of a few timed runs) whose outputs are
identical to the default setting's.  The result is written to tuning.json beside this module; apply() sets a
class's knobs for the process.  Classes: "lidar_frame" (voxel -> SOR k=50 -> radius normals -> RANSAC on a
1.04M-point frame, two frames in flight as the benchmark runs it; ms per frame), "batch" (256 clouds: SOR k=32 ->
radius lists -> k-NN normals k=24) and "indoor" (20M points: radius lists r=0.02 cap 64 and k-NN normals k=30).

Usage (on a GPU): python -m rvpoint_amd.autotune [--out PATH] [class ...]   (PATH defaults to tuning.json here)
"""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")

DEFAULTS = {"knn_cell_div": 2.4, "seg_cells_per_point": 32.0, "queries_per_wave": 0.0,
            "work_steal": 0.0, "knn_start_cells": 1.0, "normals_per_wave": 0.0}

# candidate values per class, in the order the descent visits the knobs
SPACE = {
    "lidar_frame": {"knn_cell_div": [1.6, 2.0, 2.4, 2.8, 3.2], "knn_start_cells": [1, 2, 3],
                    "queries_per_wave": [0, 1, 2, 4], "normals_per_wave": [0, 2, 3, 4, 6],
                    "work_steal": [0, 2], "seg_cells_per_point": [8, 16, 32, 64]},
    "batch": {"knn_cell_div": [1.6, 2.0, 2.4, 2.8, 3.2], "knn_start_cells": [1, 2, 3],
              "queries_per_wave": [0, 4, 8, 16, 32],
              "work_steal": [0, 2], "seg_cells_per_point": [8, 16, 32, 64]},
    "indoor": {"knn_cell_div": [1.6, 2.0, 2.4, 2.8, 3.2], "knn_start_cells": [1, 2, 3],
               "queries_per_wave": [0, 16, 32, 64]},
}


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def apply(cls, table=None):
    """Set the knobs of workload class `cls` from the table (the library defaults for knobs it does not list);
    returns the knobs set."""
    table = load_table() if table is None else table
    knobs = dict(DEFAULTS)
    knobs.update(table.get("classes", {}).get(cls, {}).get("knobs", {}))
    for k, v in knobs.items():
        _lib.tune(k, v)
    return knobs


# ---------------------------------------------------------------------------------------------------------
# workloads: each returns (run, outputs) closures; run() does one timed pass, outputs() the results to compare
# ---------------------------------------------------------------------------------------------------------
def _frame_workload():
    """The frame as the bench's headline runs it: two frames in flight (ConcurrentFrames), 20 frames per pass.
    The pipelines' graphs fix the knobs at capture, so a pass under other knobs captures anew (in the untimed
    first pass of each setting)."""
    from . import scenes
    from .pipeline import ConcurrentFrames
    frame = scenes.lidar_frame()
    state = {"knobs": None, "cf": None}

    def current():
        if state["knobs"] != {k: _lib.tune(k) for k in DEFAULTS}:
            if state["cf"] is not None:
                state["cf"].close()
            state["cf"] = ConcurrentFrames(frame, n=2, keep_means=True)
            state["knobs"] = {k: _lib.tune(k) for k in DEFAULTS}
        return state["cf"]

    def run():
        current().run(10)
    run.frames = 20  # timed per frame

    def outputs():
        r = current().pipes[0].results()
        return [r["voxels"], r["means"], r["kept"], r["normals"], r["model"], np.array([r["inliers"], r["iterations"]])]
    return run, outputs


def _batch_workload():
    from . import scenes
    from .pipeline import DeviceBatchPipeline
    clouds, scales = scenes.batch_clouds(with_scales=True)
    bp = DeviceBatchPipeline(clouds, scales)
    del clouds
    rows = 250_000  # list rows compared (the full lists are 11 GB); counts and normals compared whole

    def run():
        bp.step()

    def outputs():
        kept = bp.kept_total()
        r = min(rows, kept)
        return [bp.means.numpy(), bp.sor_off.numpy(), bp.cnt.numpy()[:kept], bp.idx.numpy(count=r * bp.CAP)[:r],
                bp.d2.numpy(count=r * bp.CAP)[:r], bp.nx.numpy()[:kept], bp.ny.numpy()[:kept], bp.nz.numpy()[:kept]]
    return run, outputs


def _indoor_workload():
    from . import scenes
    from ._lib import call
    from .device import DeviceArray, DeviceCloud
    import ctypes as C
    pts = scenes.indoor_scan()
    n = len(pts)
    dev = DeviceCloud(pts)
    del pts
    h = C.c_void_p()
    call("rvp_hash_create", C.byref(h))
    call("rvp_hash_build", h.value, *dev.ptrs(), n, 0.02, None)
    idx, d2, cnt = DeviceArray((n, 64), np.int32), DeviceArray((n, 64)), DeviceArray(n, np.int32)
    nx, ny, nz = DeviceArray(n), DeviceArray(n), DeviceArray(n)

    def run():
        call("rvp_radius_search", h.value, None, None, None, n, 0.02, 64, idx.ptr, d2.ptr, cnt.ptr, None, None)
        call("rvp_normals_knn", *dev.ptrs(), n, 30, 0.0, 0.0, 0.0, nx.ptr, ny.ptr, nz.ptr, None)

    def outputs():
        r = 1_000_000  # list rows compared (the full lists are 10 GB); counts and normals compared whole
        return [cnt.numpy(), idx.numpy(count=r * 64)[:r], d2.numpy(count=r * 64)[:r], nx.numpy(), ny.numpy(),
                nz.numpy()]
    return run, outputs


WORKLOADS = {"lidar_frame": _frame_workload, "batch": _batch_workload, "indoor": _indoor_workload}


def _time(run, reps):
    from .device import synchronize
    run()
    synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        synchronize()
        ts.append((time.perf_counter() - t0) * 1e3 / getattr(run, "frames", 1))
    return float(np.median(ts))


def _same(a, b):
    return len(a) == len(b) and all(np.array_equal(u, v) for u, v in zip(a, b))


def tune(cls, reps=None, log=print):
    """Coordinate descent over SPACE[cls]; returns {"knobs", "ms", "default_ms", "sweep"}."""
    reps = reps or (7 if cls == "lidar_frame" else 3)
    run, outputs = WORKLOADS[cls]()
    apply_knobs = lambda kn: [_lib.tune(k, v) for k, v in kn.items()]  # noqa: E731
    best = dict(DEFAULTS)
    apply_knobs(best)
    run()
    ref = [np.copy(o) for o in outputs()]
    best_ms = default_ms = _time(run, reps)
    log(f"autotune {cls}: default {default_ms:.3f} ms")
    sweep = []
    for knob, values in SPACE[cls].items():
        for v in values:
            if float(v) == best[knob]:
                continue
            trial = dict(best, **{knob: float(v)})
            apply_knobs(trial)
            ms = _time(run, reps)
            same = _same(outputs(), ref)
            sweep.append({"knob": knob, "value": float(v), "ms": round(ms, 4), "same_results": bool(same)})
            log(f"autotune {cls}: {knob}={v}: {ms:.3f} ms (best {best_ms:.3f}){'' if same else ' RESULTS DIFFER'}")
            if same and ms < best_ms:
                best, best_ms = trial, ms
        apply_knobs(best)
    return {"knobs": best, "ms": round(best_ms, 4), "default_ms": round(default_ms, 4), "reps": reps,
            "sweep": sweep}


def main(argv):
    out = TABLE
    if argv[:1] == ["--out"]:
        out, argv = argv[1], argv[2:]
    classes = argv or list(WORKLOADS)
    if not _lib.gpu_available():
        print("autotune needs a GPU", file=sys.stderr)
        return 1
    table = load_table(out)
    table.setdefault("classes", {})
    for cls in classes:
        res = tune(cls)
        table["classes"][cls] = res
        with open(out, "w") as f:
            json.dump(table, f, indent=1)
        print(json.dumps({cls: {k: res[k] for k in ("knobs", "ms", "default_ms")}}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
