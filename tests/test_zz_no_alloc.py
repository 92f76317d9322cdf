"""rvp_plane_extract_noalloc (the allocation-free plane split) against rvp_plane_extract through the C ABI: same
counts and the same inlier / outlier arrays, bit for bit, on two calls in a row (the second smaller, so the
scratch holds the first call's words), one of them with a NULL output; sizes above the fixed scratch are
rejected with RVP_E_UNSUPPORTED before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = pytest.importorskip("rvpoint_amd")


def split(name, dev, n, model, thr, want_in=True, want_out=True):
    from rvpoint_amd._lib import call
    din = R.DeviceArray((max(n, 1), 3)) if want_in else None
    dout = R.DeviceArray((max(n, 1), 3)) if want_out else None
    ni, no = C.c_size_t(), C.c_size_t()
    m = np.ascontiguousarray(model, np.float32)
    call(name, *dev.ptrs(), n, m.ctypes.data_as(C.POINTER(C.c_float)), float(thr),
         din.ptr if din else None, C.byref(ni), dout.ptr if dout else None, C.byref(no), None)
    return (ni.value, no.value, din.numpy()[:ni.value] if din else None, dout.numpy()[:no.value] if dout else None)


def check_same(p, model, thr, want_in=True, want_out=True):
    dev = R.DeviceCloud(p)
    a = split("rvp_plane_extract", dev, len(p), model, thr, want_in, want_out)
    b = split("rvp_plane_extract_noalloc", dev, len(p), model, thr, want_in, want_out)
    assert a[0] == b[0] and a[1] == b[1] and a[0] + a[1] == len(p)
    for u, v in zip(a[2:], b[2:]):
        assert (u is None and v is None) or np.array_equal(u, v)


def test_noalloc_plane_split_matches_the_allocating_entry_point():
    rng = np.random.default_rng(7)
    big = np.concatenate([np.c_[rng.uniform(-10, 10, (3_000_000, 2)), rng.normal(10, 0.02, 3_000_000)],
                          rng.uniform(-20, 20, (1_000_003, 3))]).astype(np.float32)
    model = np.array([0.0, 0.0, 1.0, -10.0], np.float32)
    check_same(big, model, 0.03)                       # first call: 4,000,003 points (1,954 tiles)
    small = big[rng.permutation(len(big))[:20_011]]
    check_same(small, model, 0.03)                     # second call, smaller: stale words of the first remain
    check_same(small[:777], model, 0.5, want_in=False)  # outliers only
    check_same(small[:1], model, 0.03)


def test_noalloc_plane_split_rejects_oversized_input():
    from rvpoint_amd import _lib
    L = _lib.lib()
    m = (C.c_float * 4)(0.0, 0.0, 1.0, 0.0)
    ni, no = C.c_size_t(), C.c_size_t()
    p = C.c_void_p(256)  # never dereferenced: the size check comes first
    st = L.rvp_plane_extract_noalloc(p, p, p, (1 << 27) + 1, m, C.c_float(0.1), p, C.byref(ni), p, C.byref(no),
                                     None)
    assert st == _lib.RVP_E_UNSUPPORTED and b"scratch" in L.rvp_last_error()
