"""Device memory owned by the HIP library, and the array type that
`return_device=True` hands back.

DeviceArray keeps the CuPy idiom the reference's callers use
(`.get()`, `.shape`, `.dtype`; tests/test_lib_gpu.py:30-33) without CuPy, and
exports DLPack (device type kDLROCM) so torch and friends can take the result
without a copy.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from ._capi import call, lib

KDL_ROCM = 10


class DevicePool:
    """Caching allocator over bl_malloc. Freed blocks go back to a free list
    keyed by their rounded size and are handed out again instead of calling
    hipFree, which synchronises the whole device and would serialise work
    running on other streams. Sizes round up to 1/8-power-of-two steps above
    1 MiB (exact 256-byte multiples below)."""

    def __init__(self):
        self._free = {}
        self._lock = threading.RLock()   # __del__ may run inside give_back via GC
        self.cached = 0
        self.live = 0

    @staticmethod
    def round_size(n: int) -> int:
        n = max(int(n), 256)
        if n <= 1 << 20:
            return (n + 255) & ~255
        step = 1 << (n.bit_length() - 4)
        return (n + step - 1) // step * step

    def alloc(self, nbytes: int):
        cap = self.round_size(nbytes)
        with self._lock:
            lst = self._free.get(cap)
            if lst:
                ptr = lst.pop()
                self.cached -= cap
                self.live += cap
                return ptr, cap
        p = ctypes.c_void_p()
        try:
            call("bl_malloc", p, cap)
        except MemoryError:
            self.release_cached()
            call("bl_malloc", p, cap)
        with self._lock:
            self.live += cap
        return p.value, cap

    def give_back(self, ptr: int, cap: int):
        with self._lock:
            self._free.setdefault(cap, []).append(ptr)
            self.cached += cap
            self.live -= cap

    def release_cached(self):
        """Return every cached block to the driver (synchronises the device)."""
        with self._lock:
            blocks = [(p, c) for c, lst in self._free.items() for p in lst]
            self._free.clear()
            self.cached = 0
        for p, _ in blocks:
            lib().bl_free(p)


POOL = DevicePool()


class DeviceBuffer:
    """One device allocation from the caching pool, returned to it with the
    object (or by free())."""

    __slots__ = ("ptr", "nbytes", "cap", "__weakref__")

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        if self.nbytes == 0:
            self.ptr, self.cap = 0, 0
        else:
            self.ptr, self.cap = POOL.alloc(self.nbytes)

    def free(self):
        if self.ptr:
            p, self.ptr = self.ptr, 0
            POOL.give_back(p, self.cap)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, host: np.ndarray, stream=None, offset: int = 0):
        """Copy host data in, ordered on `stream`. Without a stream the copy
        has landed when this returns: the null stream is not ordered with the
        non-blocking streams the package creates, so a kernel launched next on
        one of them must not race an in-flight copy on it."""
        host = np.ascontiguousarray(host)
        if host.nbytes + offset > self.nbytes:
            raise ValueError("upload larger than the buffer")
        call("bl_memcpy_htod", self.ptr + offset, host.ctypes.data, host.nbytes, stream)
        if stream is None and host.nbytes:
            call("bl_stream_sync", None)

    @classmethod
    def from_host(cls, host: np.ndarray, stream=None) -> "DeviceBuffer":
        host = np.ascontiguousarray(host)
        b = cls(max(host.nbytes, 1))
        b.upload(host, stream)
        return b


class DeviceArray:
    """A typed view of device memory. `.get()` copies to a numpy array."""

    def __init__(self, shape, dtype, buffer: DeviceBuffer | None = None, offset: int = 0):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._buf = buffer if buffer is not None else DeviceBuffer(nbytes)
        self._off = int(offset)
        if self._off + nbytes > self._buf.nbytes:
            raise ValueError("view exceeds its buffer")

    @classmethod
    def empty(cls, shape, dtype) -> "DeviceArray":
        return cls(shape, dtype)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def ptr(self) -> int:
        """Device address of element 0 (0 for an empty array)."""
        return self._buf.ptr + self._off if self._buf.ptr else 0

    def __len__(self):
        return self.shape[0] if self.shape else 0

    def reshape(self, *shape) -> "DeviceArray":
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        if int(np.prod(shape, dtype=np.int64)) != self.size:
            raise ValueError(f"cannot reshape {self.shape} into {shape}")
        return DeviceArray(shape, self.dtype, self._buf, self._off)

    def get(self, stream=None) -> np.ndarray:
        """Copy to a new numpy array (waits for the copy)."""
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            call("bl_memcpy_dtoh", out.ctypes.data, self.ptr, self.nbytes, stream)
            call("bl_stream_sync", stream)
        return out

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"

    # ------------------------------------------------------------ DLPack export
    def __dlpack_device__(self):
        return (KDL_ROCM, 0)

    def __dlpack__(self, stream=None, **_):
        return _dlpack_capsule(self)


# DLPack v0.x structures (dlpack.h), enough to hand a contiguous array over.
class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int), ("device_id", ctypes.c_int)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


_DELETER = ctypes.CFUNCTYPE(None, ctypes.POINTER(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", _DELETER)]

_live = {}          # id -> (managed tensor, shape array, owner) kept until the consumer deletes


@_DELETER
def _dl_deleter(ptr):
    _live.pop(ctypes.addressof(ptr.contents), None)


def _dlpack_capsule(arr: DeviceArray):
    code = {"i": 0, "u": 1}[arr.dtype.kind]
    shape = (ctypes.c_int64 * max(arr.ndim, 1))(*arr.shape)
    mt = _DLManagedTensor()
    mt.dl_tensor.data = arr.ptr
    mt.dl_tensor.device = _DLDevice(KDL_ROCM, 0)
    mt.dl_tensor.ndim = arr.ndim
    mt.dl_tensor.dtype = _DLDataType(code, arr.dtype.itemsize * 8, 1)
    mt.dl_tensor.shape = ctypes.cast(shape, ctypes.POINTER(ctypes.c_int64))
    mt.dl_tensor.strides = None
    mt.dl_tensor.byte_offset = 0
    mt.deleter = _dl_deleter
    _live[ctypes.addressof(mt)] = (mt, shape, arr)
    new = ctypes.pythonapi.PyCapsule_New
    new.restype = ctypes.py_object
    new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    return new(ctypes.addressof(mt), b"dltensor", None)
