"""The C++ drop-in layer (include/rvv_pcl.h + librvpoint_amd.so).

CPU: the reference's own unit-test sources compile unchanged against include/ (when the reference tree is
mounted; syntax check only).  GPU: tests/cpp/unit_checks runs the reference unit tests' checks through the
drop-in API on the same std::mt19937(42) inputs."""
import os
import subprocess

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
