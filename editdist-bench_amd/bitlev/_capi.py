"""ctypes binding of libbitlev_hip.so (C ABI declared in include/bitlev_hip.h).

The library is loaded from this package directory. A missing or unloadable
library raises BackendUnavailable: there is no CPU fallback anywhere in the
backend, so nothing can silently compute distances without the GPU.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbitlev_hip.so")

BL_OK, BL_EALLOC, BL_EINVAL, BL_EHIP, BL_ENODEV = 0, -1, -2, -3, -4

_c = ctypes
_vp, _i64, _i32, _int, _sz = _c.c_void_p, _c.c_int64, _c.c_int32, _c.c_int, _c.c_size_t

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "bl_version": [],
    "bl_last_error": [],
    "bl_device_count": [],
    "bl_set_device": [_int],
    "bl_device_info": [_c.c_char_p, _sz, _c.POINTER(_int), _c.POINTER(_int), _c.POINTER(_int)],
    "bl_malloc": [_c.POINTER(_vp), _sz],
    "bl_free": [_vp],
    "bl_host_malloc": [_c.POINTER(_vp), _sz],
    "bl_host_free": [_vp],
    "bl_memcpy_htod": [_vp, _vp, _sz, _vp],
    "bl_memcpy_dtoh": [_vp, _vp, _sz, _vp],
    "bl_memcpy_dtod": [_vp, _vp, _sz, _vp],
    "bl_memset": [_vp, _int, _sz, _vp],
    "bl_stream_create": [_c.POINTER(_vp)],
    "bl_stream_destroy": [_vp],
    "bl_stream_sync": [_vp],
    "bl_device_sync": [],
    "bl_event_create": [_c.POINTER(_vp)],
    "bl_event_destroy": [_vp],
    "bl_event_record": [_vp, _vp],
    "bl_stream_wait_event": [_vp, _vp],
    "bl_event_sync": [_vp],
    "bl_event_query": [_vp],
    "bl_event_elapsed_ms": [_vp, _vp, _c.POINTER(_c.c_float)],
    "bl_graph_begin": [_vp],
    "bl_graph_end": [_vp, _c.POINTER(_vp)],
    "bl_graph_launch": [_vp, _vp],
    "bl_graph_destroy": [_vp],
    "bl_pairs_class": [_int, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _int, _int, _int, _vp],
    "bl_stripes_layout": [_vp, _i64, _vp],
    "bl_stripes_build": [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    "bl_gather_cols": [_vp, _vp, _vp, _i64, _i64, _int, _vp],
    "bl_cdist_class": [_int, _vp, _vp, _vp, _i64, _vp, _vp, _int, _int, _vp],
    "bl_long": [_int, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _int, _int, _int, _i64, _vp],
    "bl_cdist_workspace_size": [_i64, _int],
    "bl_cdist": [_vp, _vp, _i64, _vp, _int, _vp, _int, _int, _vp, _sz, _vp],
    "bl_cdist_tuned": [_vp, _vp, _i64, _vp, _int, _vp, _int, _int, _vp, _vp, _sz, _vp, _vp],
    "bl_cdist_matches": [_vp, _vp, _i64, _vp, _int, _int, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _sz,
                         _vp, _vp],
    "bl_pairs_workspace_size": [_i64, _i64],
    "bl_pairs": [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _int, _int, _int, _vp, _sz, _vp],
    "bl_pairs_tuned": [_vp, _vp, _vp, _vp, _i64, _i64, _vp, _int, _int, _int, _vp, _vp, _sz, _vp,
                       _vp],
}
_RESTYPE = {"bl_last_error": _c.c_char_p, "bl_cdist_workspace_size": _sz,
            "bl_pairs_workspace_size": _sz, "bl_stripes_layout": _i64}


class Stripes(ctypes.Structure):
    """bl_stripes (include/bitlev_hip.h)."""
    _fields_ = [("words", _vp), ("gbase", _vp), ("len", _vp), ("n", _i64), ("counts", _vp),
                ("counts2", _vp)]


class CdistParams(ctypes.Structure):
    """bl_cdist_params (include/bitlev_hip.h)."""
    _fields_ = [("waves_per_cu", _int), ("tile", _i64), ("reserved", _int * 4)]


class PairsParams(ctypes.Structure):
    """bl_pairs_params (include/bitlev_hip.h)."""
    _fields_ = [("waves_per_cu", _int), ("block", _int), ("bin_width", _int), ("reserved", _int * 4)]


class BackendUnavailable(RuntimeError):
    """The HIP library is missing, failed to load, or no GPU is visible."""


class HipError(RuntimeError):
    """A HIP runtime failure reported by the library."""


_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library (raises BackendUnavailable if it cannot be loaded)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise BackendUnavailable(
                    f"{LIB_PATH} is not built; run `make -C editdist-bench_amd` (hipcc, gfx950)")
            try:
                L = ctypes.CDLL(LIB_PATH)
            except OSError as e:
                raise BackendUnavailable(f"cannot load {LIB_PATH}: {e}") from e
            for name, args in SIGNATURES.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = _RESTYPE.get(name, _int)
            _lib = L
    return _lib


def check(rc: int, what: str = "") -> int:
    """Map a C ABI status to the reference's exception types (bitlev/gpu.py,
    csrc/edpy.c:143-144, :185-186): -1 MemoryError, -2 ValueError, else HipError."""
    if rc >= 0:
        return rc
    msg = (lib().bl_last_error() or b"").decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == BL_EALLOC:
        raise MemoryError(text)
    if rc == BL_EINVAL:
        raise ValueError(text)
    if rc == BL_ENODEV:
        raise BackendUnavailable(text)
    raise HipError(text)


def call(name: str, *args) -> int:
    return check(getattr(lib(), name)(*args), name)
