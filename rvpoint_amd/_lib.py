"""ctypes binding of librvp_hip.so (the C ABI declared in include/rvp.h).

The library is loaded from rvpoint_amd/lib/ (built in-tree by rvpoint_amd.build).  There is no CPU
fallback: when the library or a GPU is missing, every call raises RvpError.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "librvp_hip.so")

RVP_OK, RVP_E_ARG, RVP_E_OOM, RVP_E_DEVICE, RVP_E_STATE, RVP_E_UNSUPPORTED = 0, -1, -2, -3, -4, -5


class RvpError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"rvp status {status}: {msg}")
        self.status = status


vp = C.c_void_p
sz = C.c_size_t
fl = C.c_float
i32 = C.c_int
pvp = C.POINTER(C.c_void_p)
psz = C.POINTER(C.c_size_t)
pi32 = C.POINTER(C.c_int)
pfl = C.POINTER(C.c_float)
pi64 = C.POINTER(C.c_int64)

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "rvp_last_error": [],
    "rvp_version": [],
    "rvp_device_count": [pi32],
    "rvp_set_device": [i32],
    "rvp_malloc": [pvp, sz],
    "rvp_free": [vp],
    "rvp_memcpy_h2d": [vp, vp, sz, vp],
    "rvp_memcpy_d2h": [vp, vp, sz, vp],
    "rvp_memcpy_d2d": [vp, vp, sz, vp],
    "rvp_memset": [vp, i32, sz, vp],
    "rvp_stream_create": [pvp],
    "rvp_stream_destroy": [vp],
    "rvp_stream_sync": [vp],
    "rvp_device_sync": [],
    "rvp_event_create": [pvp],
    "rvp_event_destroy": [vp],
    "rvp_event_record": [vp, vp],
    "rvp_event_elapsed": [vp, vp, pfl],
    "rvp_stream_wait_event": [vp, vp],
    "rvp_aos_to_soa": [vp, sz, vp, vp, vp, vp],
    "rvp_aos_to_soa_bounded": [vp, sz, vp, vp, vp, vp, vp],
    "rvp_graph_capture_begin": [vp],
    "rvp_graph_capture_end": [vp, pvp],
    "rvp_graph_launch": [vp, vp],
    "rvp_graph_destroy": [vp],
    "rvp_profile_enable": [i32],
    "rvp_profile_collect": [C.c_char_p, sz, psz],
    "rvp_profile_filter": [C.c_char_p],
    "rvp_cache_trim": [],
    "rvp_exclusive_scan_u32": [vp, vp, sz, vp, vp],
    "rvp_scan_wait_errors": [C.POINTER(C.c_ulonglong)],
    "rvp_sort_pairs_u32": [vp, vp, vp, vp, sz, C.c_int, vp, vp, C.POINTER(C.c_int), vp],
    "rvp_sort_pairs_u64": [vp, vp, vp, vp, sz, C.c_int, vp, vp, C.POINTER(C.c_int), vp],
    "rvp_tune_set": [C.c_char_p, C.c_double],
    "rvp_tune_get": [C.c_char_p, C.POINTER(C.c_double)],
    "rvp_dist_sq": [vp, vp, vp, sz, fl, fl, fl, vp, vp],
    "rvp_inds_in_radius": [vp, vp, vp, vp, sz, fl, fl, fl, fl, vp, vp, psz, vp],
    "rvp_hash_create": [pvp],
    "rvp_hash_destroy": [vp],
    "rvp_hash_build": [vp, vp, vp, vp, sz, fl, vp],
    "rvp_hash_bbox": [vp, pfl],
    "rvp_hash_info": [vp, psz, psz, pi32, pfl],
    "rvp_hash_cell_coords": [vp, vp, vp, vp],
    "rvp_radius_search": [vp, vp, vp, vp, sz, fl, i32, vp, vp, vp, vp, vp],
    "rvp_radius_count": [vp, vp, vp, vp, sz, fl, vp, vp],
    "rvp_radius_fill": [vp, vp, vp, vp, sz, fl, vp, vp, vp, vp],
    "rvp_exclusive_scan": [vp, sz, vp, pi64, vp],
    "rvp_radius_bruteforce": [vp, vp, vp, sz, vp, vp, vp, sz, fl, i32, vp, vp, vp, vp],
    "rvp_knn": [vp, vp, vp, vp, sz, i32, vp, vp, vp],
    "rvp_sor_exact": [vp, vp, vp, sz, i32, fl, vp, psz, vp, vp, vp],
    "rvp_sor_radius": [vp, i32, fl, fl, vp, psz, vp, vp, vp],
    "rvp_seq_sum": [vp, sz, i32, vp, vp, vp],
    "rvp_normals_knn": [vp, vp, vp, sz, i32, fl, fl, fl, vp, vp, vp, vp],
    "rvp_normals_knn_centroids": [vp, vp, vp, sz, i32, fl, fl, fl, vp, vp, vp, vp, vp, vp, vp],
    "rvp_normals_radius": [vp, fl, i32, fl, fl, fl, vp, vp, vp, vp],
    "rvp_normals_fallback_count": [psz],
    "rvp_voxel": [vp, vp, vp, sz, fl, vp, psz, vp],
    "rvp_ransac_plane": [vp, vp, vp, sz, fl, i32, fl, fl, pfl, pi32, pi32, vp],
    "rvp_plane_counts": [vp, vp, vp, sz, vp, i32, fl, vp, vp],
    "rvp_plane_extract": [vp, vp, vp, sz, pfl, fl, vp, psz, vp, psz, vp],
    "rvp_plane_extract_noalloc": [vp, vp, vp, sz, pfl, fl, vp, psz, vp, psz, vp],
    "rvp_ransac_plane_extract": [vp, vp, vp, sz, fl, i32, fl, fl, pfl, pi32, pi32, vp, psz, vp, psz, vp],
    "rvp_batch_sor_exact": [vp, vp, vp, vp, i32, sz, i32, fl, vp, vp, vp, vp, vp],
    "rvp_batch_radius_search": [vp, vp, vp, vp, i32, sz, vp, vp, i32, vp, vp, vp, vp],
    "rvp_batch_normals_knn": [vp, vp, vp, vp, i32, sz, i32, fl, fl, fl, vp, vp, vp, vp],
    "rvp_batch_normals_knn_lists": [vp, vp, vp, vp, i32, sz, i32, vp, vp, vp, i32, fl, fl, fl, vp, vp, vp, vp],
    "rvp_batch_normals_radius": [vp, vp, vp, vp, i32, sz, vp, vp, i32, fl, fl, fl, vp, vp, vp, vp],
    "rvp_batch_voxel": [vp, vp, vp, vp, i32, sz, vp, vp, vp, vp, vp],
    "rvp_batch_ransac_plane": [vp, vp, vp, vp, i32, sz, fl, i32, fl, fl, pfl, pi32, pi32, vp],
    "rvp_euclidean_cluster": [vp, fl, i32, i32, vp, vp, psz, vp, vp],
    "rvp_pcd_info": [C.c_char_p, psz, pi32],
    "rvp_pcd_read": [C.c_char_p, pfl, pfl, pfl, sz, psz],
    "rvp_pcd_read_device": [C.c_char_p, vp, vp, vp, sz, psz, vp],
    "rvp_pcd_write": [C.c_char_p, pfl, pfl, pfl, sz, i32],
    "rvp_cluster_errors": [C.POINTER(C.c_ulonglong)],
    "rvp_server_create": [vp, vp, i32, sz, sz],
    "rvp_server_slots": [vp, pi32],
    "rvp_server_submit": [vp, pfl, sz, vp],
    "rvp_server_wait": [vp, C.c_uint64, vp],
    "rvp_result_free": [vp],
    "rvp_server_destroy": [vp],
}
_RESTYPE = {"rvp_last_error": C.c_char_p, "rvp_version": C.c_char_p}

_lib = None


def lib():
    """Load librvp_hip.so once and declare the C ABI."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RvpError(RVP_E_STATE, f"{LIB_PATH} not built (python -m rvpoint_amd.build)")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


def call(name, *args):
    st = getattr(lib(), name)(*args)
    if st != RVP_OK:
        raise RvpError(st, lib().rvp_last_error().decode())
    return st


def tune(name, value=None):
    """Read (value None) or set a process-wide tuning knob; returns the value in effect."""
    if value is not None:
        call("rvp_tune_set", name.encode(), float(value))
    v = C.c_double()
    call("rvp_tune_get", name.encode(), C.byref(v))
    return v.value


def profile_enable(on=True, kernel=None):
    """Time kernel launches with HIP events on their own streams (optionally only `kernel`)."""
    call("rvp_profile_filter", kernel.encode() if kernel else None)
    call("rvp_profile_enable", 1 if on else 0)


def profile_collect():
    """{kernel: (launches, total_ms)} of the launches recorded since the last collect."""
    need = sz()
    call("rvp_profile_collect", None, 0, C.byref(need))
    buf = C.create_string_buffer(max(need.value, 1))
    call("rvp_profile_collect", buf, len(buf), C.byref(need))
    out = {}
    for line in buf.value.decode().splitlines():
        name, n, ms = line.split()
        c, t = out.get(name, (0, 0.0))
        out[name] = (c + int(n), t + float(ms))
    return out


def gpu_available() -> bool:
    try:
        n = C.c_int(0)
        return lib().rvp_device_count(C.byref(n)) == RVP_OK and n.value > 0
    except (RvpError, OSError):
        return False
