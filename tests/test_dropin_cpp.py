"""The C++ drop-in layer (include/rvv_pcl.h + librvpoint_amd.so).

CPU: the reference's own unit-test sources compile unchanged against include/ (when the reference tree is
mounted; syntax check only).  GPU: tests/cpp/unit_checks runs the reference unit tests' checks through the
drop-in API on the same std::mt19937(42) inputs."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TESTS = "/root/reference/tests/unit"
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "unit_checks")


@pytest.mark.skipif(not os.path.isdir(REF_TESTS), reason="reference tree not mounted")
@pytest.mark.parametrize("name", ["test_radius.cpp", "test_sor.cpp", "test_voxel_grid.cpp", "test_ransac.cpp",
                                  "test_euclidean_clustering.cpp", "test_octree.cpp", "test_normal.cpp"])
def test_reference_unit_tests_compile_against_dropin(name):
    inc = os.path.join(ROOT, "include")
    dirs = [inc] + [os.path.join(inc, d) for d in ("core", "search", "filters", "features", "segmentation")]
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-deprecated-declarations"]
                       + [a for d in dirs for a in ("-I", d)] + [os.path.join(REF_TESTS, name)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


REF_UNIT = os.path.join(ROOT, "oracle", "_ref", "unit")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["test_radius", "test_sor", "test_voxel_grid", "test_ransac",
                                  "test_euclidean_clustering", "test_octree", "test_normal"])
def test_reference_unit_tests_pass_on_the_gpu(name):
    """The reference's own unit-test programs, built unchanged against the drop-in (oracle/_ref/unit)."""
    exe = os.path.join(REF_UNIT, name)
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/unit not built (needs the reference tree at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]


@pytest.mark.gpu
def test_unit_checks_program():
    assert os.path.exists(EXE), "tests/cpp/_build/unit_checks not built (python -m rvpoint_amd.build)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_reference_pipeline_walkthrough_on_the_gpu():
    """tests/integration/test_pipeline_walkthrough.cpp, built unchanged against the drop-in, on the reference's
    bunny.pcd (a fixture here) with its 1,000 synthetic ground points: voxel 0.01 -> RANSAC 1000 (thr 0.005) ->
    split -> sor_rvv k=50 -> Octree normals r=0.03 -> radius search.  Its own checks must pass, and every count it
    prints must equal the oracle's on the same input (tests/cpp/walkthrough_input writes that input)."""
    import oracle as O
    exe = os.path.join(REF_UNIT, "test_pipeline_walkthrough")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/unit not built (needs the reference tree at build time)")
    pcd = os.path.join(ROOT, "tests", "golden", "bunny.pcd")
    with tempfile.TemporaryDirectory() as d:
        os.mkdir(os.path.join(d, "results"))  # where the program saves its ground and object clouds
        r = subprocess.run([exe, pcd], capture_output=True, text=True, timeout=300, cwd=d)
        print(r.stdout[-3000:])
        assert r.returncode == 0 and "[SUCCESS]" in r.stdout and "[FAIL]" not in r.stdout + r.stderr, \
            (r.stdout + r.stderr)[-3000:]
        inp = os.path.join(d, "input.bin")
        helper = os.path.join(ROOT, "tests", "cpp", "_build", "walkthrough_input")
        assert subprocess.run([helper, pcd, inp], timeout=60).returncode == 0
        raw = open(inp, "rb").read()
        saved = sorted(os.listdir(os.path.join(d, "results")))
    pts = np.frombuffer(raw[8:], np.float32).reshape(-1, 3)
    assert len(pts) == int(np.frombuffer(raw[:8], np.uint64)[0]) == 397 + 1000
    out = r.stdout
    num = lambda pat: int(re.search(pat, out).group(1))  # noqa: E731
    vox = O.voxel_sc(pts, 0.01)
    assert num(r"Filtered count: (\d+)") == len(vox)
    x, y, z = (np.ascontiguousarray(vox[:, i]) for i in range(3))
    best, model, _ = O.ransac_plane(x, y, z, 0.005, 1000)
    assert num(r"RANSAC found (\d+) plane inliers") == best
    ins, outs = O.plane_extract(x, y, z, model, 0.005)
    assert num(r"Extracted: (\d+) plane points") == len(ins)
    assert num(r"plane points \(ground\), (\d+) object points") == len(outs)
    kept, _ = O.sor_sc(outs, 50, 1.0)
    assert num(r"SOR Filtered count: (\d+)") == len(kept)
    assert len(saved) == 2  # ground and voxelised object clouds written by savePCD
