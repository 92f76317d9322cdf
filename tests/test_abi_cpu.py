"""CPU-side checks of the C ABI library: it loads, exports exactly what include/rvp.h declares, and
reports errors (not crashes) without a GPU.  No compute calls."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "rvpoint_amd", "lib", "librvp_hip.so")


def declared():
    txt = open(os.path.join(ROOT, "include", "rvp.h")).read()
    return set(re.findall(r"\b(rvp_[a-z0-9_]+)\s*\(", txt))


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(SO):
        from rvpoint_amd import build
        build.build()
    return SO


def test_exports_match_header(built):
    out = subprocess.run(["nm", "-D", "--defined-only", built], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T rvp_" in l}
    assert exported == declared()


def test_python_binding_declares_every_symbol(built):
    from rvpoint_amd import _lib
    assert set(_lib.exported_symbols()) == declared()
    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name)


def test_library_is_gfx950(built):
    blob = open(built, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_errors_without_device(built):
    import ctypes as C

    from rvpoint_amd import _lib
    L = _lib.lib()
    assert L.rvp_version().decode().startswith("rvpoint-amd")
    # bad arguments are rejected before any device work
    assert L.rvp_hash_create(None) == _lib.RVP_E_ARG
    assert L.rvp_radius_search(None, None, None, None, 0, C.c_float(1.0), 4, None, None, None, None, None) \
        == _lib.RVP_E_ARG
    assert b"null" in L.rvp_last_error()
    # batched entry points: empty batch, k out of range, missing buffers
    p = C.c_void_p(16)
    assert L.rvp_batch_sor_exact(p, p, p, p, 0, 8, 4, C.c_float(1.0), p, p, None, None, None) == _lib.RVP_E_ARG
    assert L.rvp_batch_sor_exact(p, p, p, p, 1, 8, 64, C.c_float(1.0), p, p, None, None, None) \
        == _lib.RVP_E_UNSUPPORTED
    assert L.rvp_batch_radius_search(p, p, p, p, 1, 8, None, None, 8, None, None, p, None) == _lib.RVP_E_ARG
    assert L.rvp_batch_normals_knn(p, p, p, p, 1, 8, 0, C.c_float(0), C.c_float(0), C.c_float(0), p, p, p,
                                   None) == _lib.RVP_E_UNSUPPORTED
    # normals from radius lists: lists required, k and the row stride in range
    f0 = C.c_float(0)
    assert L.rvp_batch_normals_knn_lists(p, p, p, p, 1, 8, 24, None, None, p, 64, f0, f0, f0, p, p, p,
                                         None) == _lib.RVP_E_ARG
    assert L.rvp_batch_normals_knn_lists(p, p, p, p, 1, 8, 64, p, None, p, 64, f0, f0, f0, p, p, p,
                                         None) == _lib.RVP_E_UNSUPPORTED
    assert L.rvp_batch_normals_knn_lists(p, p, p, p, 1, 8, 24, p, None, p, 65, f0, f0, f0, p, p, p,
                                         None) == _lib.RVP_E_ARG
    assert L.rvp_batch_normals_knn_lists(p, p, p, p, 0, 8, 24, p, p, p, 64, f0, f0, f0, p, p, p,
                                         None) == _lib.RVP_E_ARG
    # server: slots and sizes are checked first
    from rvpoint_amd.serve import FrameParams
    h = C.c_void_p()
    assert L.rvp_server_create(C.byref(h), C.byref(FrameParams.lidar()), 0, 1 << 20, 1 << 30) == _lib.RVP_E_ARG
    assert L.rvp_server_create(C.byref(h), C.byref(FrameParams.lidar()), 1, 16, 1 << 30) == _lib.RVP_E_ARG


def test_tuning_knobs_set_get_and_ranges(built):
    from rvpoint_amd import _lib
    L = _lib.lib()
    for name, good, bad in (("knn_cell_div", 1.7, 20.0), ("seg_cells_per_point", 8.0, 0.5),
                            ("queries_per_wave", 7, 65), ("knn_start_cells", 3, 9),
                            ("normals_per_wave", 4, 65)):
        old = _lib.tune(name)
        try:
            assert _lib.tune(name, good) == good
            with pytest.raises(_lib.RvpError):
                _lib.tune(name, bad)
            assert _lib.tune(name) == good  # a rejected value leaves the knob as it was
        finally:
            _lib.tune(name, old)
    assert _lib.tune("knn_cell_div") == 2.4 and _lib.tune("queries_per_wave") == 0
    with pytest.raises(_lib.RvpError):
        _lib.tune("queries_per_wave", 2.5)
    with pytest.raises(_lib.RvpError):
        _lib.tune("knn_start_cells", 1.5)
    with pytest.raises(_lib.RvpError):
        _lib.tune("no_such_knob")
    assert L.rvp_tune_get(None, None) == _lib.RVP_E_ARG
