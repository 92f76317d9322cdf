"""Build librvp_hip.so (HIP kernels + C ABI, gfx950) and librvpoint_amd.so (the C++ rvpoint:: drop-in layer)
in-tree under rvpoint_amd/lib/.  Usage: python -m rvpoint_amd.build"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")

HIP_SOURCES = ["runtime.hip", "prims.hip", "grid.hip", "query.hip", "normals.hip", "voxel_ransac.hip", "distance.hip", "seqsum.hip", "batch.hip", "server.hip", "cluster.hip", "pcd.hip"]
CXX_SOURCES = ["rvpoint_api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RVP_OFFLOAD_ARCH", "gfx950")
# -ffp-contract=off: every fp32 op of the reference path rounds once (no FMA contraction).
# Packed fp32 VALU forms are disabled for the whole library.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
             "-I", os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    """Compile every source (always; this is the 'does it build' check) and link the two libraries."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    hip_so = os.path.join(LIB, "librvp_hip.so")
    tmp = hip_so + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, hip_so)
    api_so = os.path.join(LIB, "librvpoint_amd.so")
    srcs = [os.path.join(CSRC, s) for s in CXX_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if srcs:
        tmp = api_so + ".tmp"
        _run(["g++", *CXX_FLAGS, "-shared", *srcs, "-o", tmp, "-L", LIB, "-lrvp_hip",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, api_so)
    # C++ programs that exercise the drop-in API the way the reference's own unit tests do
    test_dir = os.path.join(ROOT, "tests", "cpp")
    if os.path.isdir(test_dir) and os.path.exists(api_so):
        out_dir = os.path.join(test_dir, "_build")
        os.makedirs(out_dir, exist_ok=True)
        for src in sorted(os.listdir(test_dir)):
            if src.endswith(".cpp"):
                exe = os.path.join(out_dir, src[:-4])
                _run(["g++", *CXX_FLAGS, os.path.join(test_dir, src), "-o", exe, "-L", LIB, "-lrvpoint_amd",
                      "-lrvp_hip", "-Wl,-rpath,$ORIGIN/../../../rvpoint_amd/lib"])
    if verbose:
        print("built", hip_so, api_so)
    return hip_so


if __name__ == "__main__":
    build(verbose=True)
