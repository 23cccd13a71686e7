"""Ingest pipeline: pairwise distances of Python str/bytes lists with packing,
host-to-device copies and kernels overlapped.

The reference packs the whole batch on one thread into a new bytes object,
copies it from pageable memory, and only then runs its kernels
(bitlev/gpu.py:79-103, csrc/edpy.c:200-224); for a large batch of short
strings the packing alone takes several hundred times as long as the
kernels. Here the pairs are cut into chunks. A thread pool packs each chunk
straight into pinned host buffers; the packer reads the list's objects in
place with the GIL released (prefetching their scattered headers), so several
chunks fill at once. The calling thread copies a packed chunk to the device on a copy
stream, runs the batched pipeline (bl_pairs) on a compute stream once the
copy's event has fired, and copies the results back into pinned memory: the
packing of later chunks, the transfers and the kernels all overlap. A chunk's
buffers are reused only after its results are collected. Every pair is
independent, so the results are those of one whole-batch call.
"""
from __future__ import annotations

import concurrent.futures
import ctypes
import os
import threading
from types import SimpleNamespace

import numpy as np

from . import _edpy, hip
from ._capi import call
from ._common import check_dtype
from .device import DeviceBuffer


def effective_threads() -> int:
    """Host threads to pack with: the CPUs this process may run on, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


_pool_lock = threading.Lock()
_pool_obj = None


def _pool() -> concurrent.futures.ThreadPoolExecutor:
    """The packing threads shared by every call of this module."""
    global _pool_obj
    with _pool_lock:
        if _pool_obj is None:
            _pool_obj = concurrent.futures.ThreadPoolExecutor(
                max_workers=effective_threads(), thread_name_prefix="bitlev-pack")
        return _pool_obj


def pack(seq, threads: int | None = None, alloc=None):
    """(uint8 buffer, int64 offsets[n + 1]) of a list/tuple of str/bytes, the
    layout of _edpy.pack, packed by the thread pool in two passes: every
    thread measures a range of items, then copies it to its place in the
    buffer. Items are read in place without the GIL (the sequence must not be
    changed during the call). alloc(nbytes) -> uint8 array of at least nbytes
    supplies the buffer (e.g. pinned memory); None allocates one."""
    if not isinstance(seq, (list, tuple)):
        seq = list(seq)
    n = len(seq)
    T = threads or effective_threads()
    step = max(4096, -(-n // (4 * T)))
    ranges = [(lo, min(lo + step, n)) for lo in range(0, n, step)]
    empty = np.empty(0, np.uint8)

    def measure(r):
        lo, hi = r
        return _edpy.pack_range(seq, lo, hi, empty, np.empty(hi - lo + 1, np.int64))[0]

    pool = _pool()
    totals = list(pool.map(measure, ranges))
    base = np.zeros(len(ranges) + 1, np.int64)
    np.cumsum(totals, out=base[1:])
    need = max(int(base[-1]), 1)
    buf = np.empty(need, np.uint8) if alloc is None else alloc(need)[:need]
    off = np.empty(n + 1, np.int64)
    off[0] = 0
    if not n:
        buf[0] = 0
        return buf, off

    def fill(j):
        lo, hi = ranges[j]
        rel = np.empty(hi - lo + 1, np.int64)
        _edpy.pack_range(seq, lo, hi, buf[base[j]:base[j + 1]], rel)
        np.add(rel[1:], base[j], out=off[lo + 1:hi + 1])

    list(pool.map(fill, range(len(ranges))))
    return buf, off


class _Pinned:
    """Pinned (page-locked) host bytes, viewed as numpy."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 16)
        p = ctypes.c_void_p()
        call("bl_host_malloc", p, self.nbytes)
        self.ptr = p.value
        self.arr = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self.ptr))

    def view(self, dtype, count: int) -> np.ndarray:
        return self.arr[: count * np.dtype(dtype).itemsize].view(dtype)

    def free(self):
        if self.ptr:
            call("bl_host_free", self.ptr)
            self.ptr = None


class _Slot:
    """Host and device buffers of one chunk in flight."""

    def __init__(self, itemsize: int):
        self.itemsize = itemsize
        self.cap_n = self.cap_a = self.cap_b = 0
        self.h = {}
        self.d = {}
        ev = ctypes.c_void_p()
        call("bl_event_create", ev)
        self.copied = ev.value
        ev = ctypes.c_void_p()
        call("bl_event_create", ev)
        self.done = ev.value
        self.chunk = None                  # (lo, hi) of the chunk whose results are in flight

    def ensure(self, n: int, bytes_a: int, bytes_b: int):
        """Buffers for n pairs of the given string bytes (grown, never shrunk)."""
        if n > self.cap_n:
            self.cap_n = max(n, 2 * self.cap_n)
            for key, size in (("off_a", 8 * (self.cap_n + 1)), ("off_b", 8 * (self.cap_n + 1)),
                              ("out", self.itemsize * self.cap_n)):
                self._grow(key, size)
        if bytes_a > self.cap_a:
            self.cap_a = max(bytes_a, 2 * self.cap_a)
            self._grow("a", self.cap_a)
        if bytes_b > self.cap_b:
            self.cap_b = max(bytes_b, 2 * self.cap_b)
            self._grow("b", self.cap_b)

    def _grow(self, key, size):
        old = self.h.get(key)
        if old is not None:
            old.free()
        self.h[key] = _Pinned(size)
        self.d[key] = DeviceBuffer(max(size, 16))

    def close(self):
        for p in self.h.values():
            p.free()
        call("bl_event_destroy", self.copied)
        call("bl_event_destroy", self.done)


class _PinnedStage:
    """One pinned host buffer, grown on demand and reused, for packing a set
    right before its upload: a copy from pinned memory goes straight to the
    device, a pageable one through the driver's bounce buffers."""

    def __init__(self):
        self.lock = threading.Lock()
        self.mem = None

    def take(self, nbytes: int) -> np.ndarray:
        if self.mem is None or self.mem.nbytes < nbytes:
            if self.mem is not None:
                self.mem.free()
            self.mem = _Pinned(max(nbytes, 1 << 20))
        return self.mem.arr


_stage = _PinnedStage()


def upload_packed(strings, stream=None):
    """hip.Packed of a list/tuple of str/bytes, packed by the thread pool into
    a reused pinned buffer and uploaded from it (the stream is synchronised
    before the buffer is reused). Falls back to ordinary packing when another
    thread holds the buffer."""
    if not _stage.lock.acquire(blocking=False):
        return hip.Packed(*pack(strings), stream=stream)
    try:
        buf, off = pack(strings, alloc=_stage.take)
        P = hip.Packed(buf, off, stream=stream)
        call("bl_stream_sync", stream)
        return P
    finally:
        _stage.lock.release()


class _RetainingWorkspace(hip.Workspace):
    """A workspace that keeps the buffers it outgrows until the pipeline ends:
    the device pool is not stream-ordered, so a block given back while a
    kernel on the compute stream still reads it could be handed to a slot
    that a worker fills on the copy stream."""

    def __init__(self):
        super().__init__()
        self.retired = []

    def get(self, nbytes: int):
        if self.buf is not None and self.buf.nbytes < nbytes:
            self.retired.append(self.buf)
            self.buf = None
        return super().get(nbytes)


def _pack_chunk(slot: _Slot, a, b, lo: int, hi: int):
    """Pack pairs lo..hi-1 into the slot's pinned buffers, growing them when a
    chunk's strings do not fit. Returns (n, bytes_a, bytes_b, max_len)."""
    n = hi - lo
    slot.ensure(n, 1, 1)
    ta, ma = _edpy.pack_range(a, lo, hi, slot.h["a"].arr, slot.h["off_a"].arr)
    tb, mb = _edpy.pack_range(b, lo, hi, slot.h["b"].arr, slot.h["off_b"].arr)
    fit_a, fit_b = ta <= slot.h["a"].nbytes, tb <= slot.h["b"].nbytes
    if not (fit_a and fit_b):
        slot.ensure(n, ta, tb)
        if not fit_a:
            ta, ma = _edpy.pack_range(a, lo, hi, slot.h["a"].arr, slot.h["off_a"].arr)
        if not fit_b:
            tb, mb = _edpy.pack_range(b, lo, hi, slot.h["b"].arr, slot.h["off_b"].arr)
    return n, ta, tb, max(ma, mb)


class Stager:
    """The pipeline's pinned slots, streams, events and workspace, kept from
    one call to the next (pinning host memory costs about
    as much as the copies it speeds up). One call at a time per stager;
    pairwise() hands each concurrent caller a stager of its own."""

    def __init__(self, threads: int | None = None):
        self.threads = threads or effective_threads()   # chunks being packed at once
        self.slots = [_Slot(4) for _ in range(self.threads + 2)]
        self.streams = []
        for _ in range(2):
            s = ctypes.c_void_p()
            call("bl_stream_create", s)
            self.streams.append(s.value)
        self.ws = _RetainingWorkspace()

    def close(self):
        for s in self.streams:
            call("bl_stream_sync", s)
        for slot in self.slots:
            slot.close()
        for s in self.streams:
            call("bl_stream_destroy", s)
        self.slots, self.streams = [], []

    def run(self, a, b, k: int, dt: np.dtype, trim: bool, chunk: int, params) -> np.ndarray:
        n = len(a)
        out = np.empty(n, dt)
        if n == 0:
            return out
        T = self.threads
        copy_s, comp_s = self.streams
        bounds = [(lo, min(lo + chunk, n)) for lo in range(0, n, chunk)]
        slots = self.slots
        ws = self.ws

        def finalize(slot: _Slot):
            if slot.chunk is None:
                return
            lo, hi = slot.chunk
            call("bl_event_sync", slot.done)
            out[lo:hi] = slot.h["out"].view(dt, hi - lo)
            slot.chunk = None

        def enqueue(i: int, packed):
            slot = slots[i % len(slots)]
            m, ta, tb, max_len = packed
            h, d = slot.h, slot.d
            call("bl_memcpy_htod", d["a"].ptr, h["a"].ptr, max(ta, 1), copy_s)
            call("bl_memcpy_htod", d["b"].ptr, h["b"].ptr, max(tb, 1), copy_s)
            call("bl_memcpy_htod", d["off_a"].ptr, h["off_a"].ptr, 8 * (m + 1), copy_s)
            call("bl_memcpy_htod", d["off_b"].ptr, h["off_b"].ptr, 8 * (m + 1), copy_s)
            call("bl_event_record", slot.copied, copy_s)
            call("bl_stream_wait_event", comp_s, slot.copied)
            A = SimpleNamespace(d_buf=d["a"], d_off=d["off_a"], n=m, max_len=max_len)
            B = SimpleNamespace(d_buf=d["b"], d_off=d["off_b"], n=m, max_len=max_len)
            res = SimpleNamespace(ptr=d["out"].ptr, dtype=dt)
            hip.pairs_batched_into(A, B, res, k, trim, comp_s, ws, params)
            call("bl_memcpy_dtoh", h["out"].ptr, d["out"].ptr, dt.itemsize * m, comp_s)
            call("bl_event_record", slot.done, comp_s)
            slot.chunk = bounds[i]

        futures = {}
        try:
            submitted = 0
            for i in range(len(bounds)):
                while submitted < len(bounds) and submitted < i + T:
                    j = submitted
                    slot = slots[j % len(slots)]
                    finalize(slot)                 # its previous chunk's results, then reuse
                    lo, hi = bounds[j]
                    futures[j] = _pool().submit(_pack_chunk, slot, a, b, lo, hi)
                    submitted += 1
                enqueue(i, futures.pop(i).result())
            for slot in slots:
                finalize(slot)
        finally:
            for f in futures.values():             # packers still writing into slots
                f.cancel()
            concurrent.futures.wait(list(futures.values()))
            call("bl_stream_sync", comp_s)
            call("bl_stream_sync", copy_s)
            for slot in slots:
                slot.chunk = None
            ws.retired.clear()                     # the streams are idle now
        return out


_MAX_IDLE = 2          # stagers kept between calls (each holds its pinned slots)
_idle = []
_idle_lock = threading.Lock()


def release():
    """Free the pinned and device memory of every idle stager and the pinned
    packing buffer."""
    with _idle_lock:
        stagers, _idle[:] = list(_idle), []
    for st in stagers:
        st.close()
    with _stage.lock:
        if _stage.mem is not None:
            _stage.mem.free()
            _stage.mem = None


def pairwise(a, b, k: int = -1, dtype=np.int32, trim: bool = True, chunk: int = 1 << 17,
             threads: int | None = None, params=None) -> np.ndarray:
    """out[i] = d(a[i], b[i]) (min(d, k + 1) for k >= 0) as a numpy array of
    `dtype`, through the staged pipeline. a, b: equal-length sequences of
    str/bytes (lists and tuples are read in place, without the GIL, and must
    not be changed while the call runs).
    params: a bl_pairs launch shape (bitlev.autotune), or None for the
    untuned default (hip.default_pairs_params).
    The stager (pinned buffers sized by the largest chunk seen) is kept for
    the next call; release() frees it."""
    if not isinstance(a, (list, tuple)):
        a = list(a)
    if not isinstance(b, (list, tuple)):
        b = list(b)
    n = len(a)
    if len(b) != n:
        raise ValueError(f"length mismatch: {n} vs {len(b)}")
    dt = check_dtype(dtype)
    if params is None:
        params = hip.default_pairs_params(int(k))
    T = threads or effective_threads()
    with _idle_lock:
        st = next((s for s in _idle if s.threads == T), None)
        if st is not None:
            _idle.remove(st)
    if st is None:
        st = Stager(T)
    try:
        # at most `chunk` pairs a chunk, but enough chunks to keep every packer busy
        step = max(1024, min(int(chunk), max(8192, -(-n // (2 * T)))))
        return st.run(a, b, int(k), dt, trim, step, params)
    finally:
        with _idle_lock:
            keep = len(_idle) < _MAX_IDLE
            if keep:
                _idle.append(st)
        if not keep:
            st.close()
