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
@pytest.mark.parametrize("name", ["test_radius.cpp", "test_sor.cpp", "test_voxel_grid.cpp", "test_ransac.cpp"])
def test_reference_unit_tests_compile_against_dropin(name):
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wno-deprecated-declarations",
                        "-I", os.path.join(ROOT, "include"), os.path.join(REF_TESTS, name)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.gpu
def test_unit_checks_program():
    assert os.path.exists(EXE), "tests/cpp/_build/unit_checks not built (python -m rvpoint_amd.build)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
