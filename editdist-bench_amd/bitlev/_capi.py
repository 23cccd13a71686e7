"""Binding of libbitlev_hip.so (C ABI declared in include/bitlev_hip.h)
through the package's C extension: `bitlev._edpy` loads the library with
dlopen and exposes each entry point as a function of the same name and
argument order, called with the GIL released (csrc/edpy_hip.c).

Pointer arguments take an int address, None, or a buffer-exporting object
(ctypes instances serve as out-parameters and as the bl_stripes /
bl_*_params structures). The library is loaded from this package directory;
a missing or unloadable library raises BackendUnavailable: there is no CPU
fallback anywhere in the backend, so nothing can silently compute distances
without the GPU.
"""
from __future__ import annotations

import ctypes
import os
import threading

from . import _edpy

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbitlev_hip.so")

BL_OK, BL_EALLOC, BL_EINVAL, BL_EHIP, BL_ENODEV = 0, -1, -2, -3, -4

_c = ctypes
_vp, _i64, _int = _c.c_void_p, _c.c_int64, _c.c_int


class Stripes(ctypes.Structure):
    """bl_stripes (include/bitlev_hip.h)."""
    _fields_ = [("words", _vp), ("gbase", _vp), ("len", _vp), ("n", _i64), ("counts", _vp),
                ("counts2", _vp), ("order", _vp), ("counts3", _vp), ("buf", _vp), ("off", _vp)]


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
    """The extension module with the library bound (raises BackendUnavailable
    if it cannot be loaded)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise BackendUnavailable(
                    f"{LIB_PATH} is not built; run `make -C editdist-bench_amd` (hipcc, gfx950)")
            try:
                _edpy.hip_open(LIB_PATH)
            except OSError as e:
                raise BackendUnavailable(str(e)) from e
            _lib = _edpy
    return _lib


def symbols() -> tuple:
    """The entry points the extension binds."""
    return _edpy.hip_symbols()


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
