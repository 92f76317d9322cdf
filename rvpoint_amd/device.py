"""Device buffers and streams owned through the C ABI (no torch dependency)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib


class Stream:
    def __init__(self):
        h = C.c_void_p()
        call("rvp_stream_create", C.byref(h))
        self.handle = h.value

    def sync(self):
        call("rvp_stream_sync", self.handle)

    def close(self):
        if self.handle:
            lib().rvp_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """A HIP event: recorded on one stream, waited for by another (rvp_stream_wait_event) or timed."""

    def __init__(self):
        h = C.c_void_p()
        call("rvp_event_create", C.byref(h))
        self.handle = h.value

    def record(self, stream):
        call("rvp_event_record", self.handle, stream)

    def close(self):
        if self.handle:
            lib().rvp_event_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A contiguous device allocation with a numpy dtype and shape."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("rvp_malloc", C.byref(p), max(self.nbytes, 16))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            call("rvp_memcpy_h2d", d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes, stream)
        return d

    def numpy(self, count=None, stream=None):
        out = np.empty(self.shape, self.dtype)
        nbytes = self.nbytes if count is None else int(count) * self.dtype.itemsize
        if nbytes:
            call("rvp_memcpy_d2h", out.ctypes.data_as(C.c_void_p), self.ptr, nbytes, stream)
        return out

    def zero(self, stream=None):
        call("rvp_memset", self.ptr, 0, self.nbytes, stream)
        return self

    def free(self):
        if getattr(self, "ptr", None):
            lib().rvp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def vp(self):
        return C.c_void_p(self.ptr)


def synchronize():
    call("rvp_device_sync")


class DeviceCloud:
    """SoA point cloud resident in HBM (the reference's PointCloudSoA, src/core/point_types.h:24-29)."""

    def __init__(self, xyz, stream=None):
        xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
        self.n = len(xyz)
        self.x = DeviceArray.from_numpy(np.ascontiguousarray(xyz[:, 0]), stream)
        self.y = DeviceArray.from_numpy(np.ascontiguousarray(xyz[:, 1]), stream)
        self.z = DeviceArray.from_numpy(np.ascontiguousarray(xyz[:, 2]), stream)

    @classmethod
    def from_soa(cls, x, y, z, stream=None):
        return cls(np.stack([np.asarray(x, np.float32), np.asarray(y, np.float32), np.asarray(z, np.float32)], 1),
                   stream)

    def ptrs(self):
        return self.x.ptr, self.y.ptr, self.z.ptr
