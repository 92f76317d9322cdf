"""The autotuner's table is applied per workload class through the library's knobs (no GPU needed: the knobs
are host state).  Knobs a class's entry does not list take the library defaults."""
import pytest

from rvpoint_amd import _lib, autotune


@pytest.fixture
def restore_knobs():
    yield
    for k, v in autotune.DEFAULTS.items():
        _lib.tune(k, v)


def test_apply_sets_class_knobs(restore_knobs):
    table = {"classes": {"batch": {"knobs": {"knn_cell_div": 2.8, "queries_per_wave": 16.0, "work_steal": 2.0}},
                         "lidar_frame": {"knobs": {"knn_cell_div": 2.0}}}}
    got = autotune.apply("batch", table)
    assert got["knn_cell_div"] == 2.8 and _lib.tune("knn_cell_div") == 2.8
    assert _lib.tune("queries_per_wave") == 16.0 and _lib.tune("work_steal") == 2.0
    assert _lib.tune("seg_cells_per_point") == autotune.DEFAULTS["seg_cells_per_point"]
    autotune.apply("lidar_frame", table)  # knobs the class does not list go back to the defaults
    assert _lib.tune("knn_cell_div") == 2.0 and _lib.tune("queries_per_wave") == 0.0
    assert _lib.tune("work_steal") == 0.0
    autotune.apply("indoor", table)       # a class the table lacks: all defaults
    assert _lib.tune("knn_cell_div") == autotune.DEFAULTS["knn_cell_div"]


def test_committed_table_is_well_formed():
    table = autotune.load_table()
    for cls, entry in table.get("classes", {}).items():
        assert cls in autotune.WORKLOADS
        assert set(entry["knobs"]) <= set(autotune.DEFAULTS)
        assert entry["ms"] <= entry["default_ms"]
        for trial in entry["sweep"]:
            assert trial["knob"] in autotune.SPACE[cls]


def test_knob_ranges_rejected():
    with pytest.raises(_lib.RvpError):
        _lib.tune("work_steal", 3)
    with pytest.raises(_lib.RvpError):
        _lib.tune("no_such_knob", 1)
