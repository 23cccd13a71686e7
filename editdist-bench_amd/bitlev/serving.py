"""Continuous-batching request loop on one GPU.

    srv = Server(slots=8, budget_bytes=32 << 30)
    srv.add_set("names", strings)                    # packed once, resident in HBM
    r1 = srv.submit_cdist(queries, "names", max_dist=2, dtype="uint8")
    r2 = srv.submit_pairwise(a, b, max_dist=None, dtype="int32")
    r1.result()        # DeviceArray (q x n), r1.latency in seconds
    srv.close()

Requests queue up; a scheduler thread admits them into free slots (one HIP
stream each) as slots and memory free up, so work of several requests runs
on the device at once. Pending requests of the same kind, resident set and
output dtype are fused into one launch of the batched kernels (a per-pair /
per-query threshold vector lets requests with different max_dist share it;
exact searches fuse only with exact searches); each request gets its own
view of the fused output, and the fused inputs are uploaded part by part,
never concatenated on the host. Eight slots by default: on the serving mix
at 75 % load, p99 fell from 22 ms with four to 14.5 ms with eight (offline
throughput 41 -> 37 G pairs/s, less fusion). Admission is by HBM
footprint: a batch is admitted only when its inputs, outputs and workspace fit
in the budget next to everything still held (live results included); a
request that can never fit is rejected with MemoryError.

Latency is arrival (submit) -> result ready on the device. Results are exact
and identical whether a request runs alone or fused with others: every pair's
distance is computed independently of its neighbours.

The reference has no serving loop; its per-call API (bitlev/gpu.py) packs and
uploads the choices on every call. This module keeps them resident.
"""
from __future__ import annotations

import collections
import ctypes
import threading
import time
import weakref

import numpy as np

from . import hip
from ._capi import call
from ._common import DTYPE_CODE, check_dtype, check_k
from .device import DeviceArray, DeviceBuffer

_ids = iter(range(1, 1 << 62))


class Request:
    """A submitted pairwise or cdist call; a future for its DeviceArray."""

    def __init__(self, kind, dtype, k, pairs, nbytes_out, payload, set_name=None, trim=True):
        self.id = next(_ids)
        self.kind = kind
        self.dtype = dtype
        self.k = k
        self.pairs = pairs                 # distances this request asks for
        self.nbytes_out = nbytes_out
        self.payload = payload             # packed host inputs
        self.set_name = set_name
        self.trim = trim
        self.t_submit = time.perf_counter()
        self.t_admit = None
        self.t_done = None
        self.batch_size = 0                # requests fused into its launch
        self._done = threading.Event()
        self._result = None
        self._error = None
        self.cancelled = False

    def done(self) -> bool:
        return self._done.is_set()

    def result(self, timeout=None) -> DeviceArray:
        if not self._done.wait(timeout):
            raise TimeoutError(f"request {self.id} not done")
        if self._error is not None:
            raise self._error
        return self._result

    def release(self):
        """Drop this request's hold on its result (the device memory of a fused
        output is freed once every request of the batch has let go of it)."""
        self._result = None

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit

    def _finish(self, result=None, error=None):
        self._result, self._error = result, error
        self.t_done = time.perf_counter()
        self._done.set()


class _Slot:
    def __init__(self, idx):
        self.idx = idx
        s = ctypes.c_void_p()
        call("bl_stream_create", s)      # default priority
        self.stream = s.value
        e = ctypes.c_void_p()
        call("bl_event_create", e)
        self.event = e.value
        self.workspace = hip.Workspace()
        self.busy = False


class Server:
    """See the module docstring. slots: concurrent batches on the device;
    budget_bytes: HBM the server may hold (inputs, outputs, live results,
    workspace; resident sets are counted too); max_batch_pairs: cap on the
    distances fused into one launch."""

    def __init__(self, slots: int = 8, budget_bytes: int = 32 << 30,
                 max_batch_pairs: int = 256 << 20, executor=None):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.budget = int(budget_bytes)
        self.max_batch_pairs = int(max_batch_pairs)
        self.sets = {}
        self._pending = collections.deque()
        self._cv = threading.Condition()
        self._in_use = 0
        self._stop = False
        self.stats = collections.Counter()
        self._exec = executor or _GpuExecutor(slots)
        self._slots = [self._exec.new_slot(i) for i in range(slots)]
        self._threads = [threading.Thread(target=self._schedule, name="bitlev-sched", daemon=True)]
        for t in self._threads:
            t.start()

    # ------------------------------------------------------------ resident sets
    def add_set(self, name: str, strings=None, packed=None):
        """Make a choice set resident (packed once, uploaded once)."""
        if packed is None:
            packed = hip.pack(strings)
        P = self._exec.resident(packed)
        with self._cv:
            if name in self.sets:
                raise ValueError(f"set {name!r} already resident")
            self._in_use += _resident_bytes(P)
            self.sets[name] = P
        return P

    def drop_set(self, name: str):
        with self._cv:
            P = self.sets.pop(name)
            self._in_use -= _resident_bytes(P)
            self._cv.notify_all()

    # ------------------------------------------------------------ submission
    def submit_pairwise(self, a, b, max_dist=None, dtype=np.int32, trim: bool = True) -> Request:
        pa = a if isinstance(a, tuple) else hip.pack(list(a))
        pb = b if isinstance(b, tuple) else hip.pack(list(b))
        n = len(pa[1]) - 1
        if len(pb[1]) - 1 != n:
            raise ValueError(f"length mismatch: {n} vs {len(pb[1]) - 1}")
        dt = check_dtype(dtype)
        req = Request("pairwise", dt, check_k(max_dist), n, n * dt.itemsize, (pa, pb), trim=trim)
        return self._enqueue(req)

    def submit_cdist(self, queries, set_name: str, max_dist=None, dtype=np.int32) -> Request:
        pq = queries if isinstance(queries, tuple) else hip.pack(list(queries))
        if set_name not in self.sets:
            raise KeyError(f"no resident set {set_name!r}")
        nq = len(pq[1]) - 1
        if nq and int(np.diff(pq[1]).max()) > hip.BATCHED_MAX_QUERY:
            raise ValueError("served cdist queries must be at most 512 bytes")
        nc = self.sets[set_name].n
        dt = check_dtype(dtype)
        req = Request("cdist", dt, check_k(max_dist), nq * nc, nq * nc * dt.itemsize, pq,
                      set_name=set_name)
        return self._enqueue(req)

    def _enqueue(self, req: Request) -> Request:
        if self._footprint([req]) > self.budget:
            self.stats["rejected"] += 1
            req._finish(error=MemoryError(
                f"request needs {self._footprint([req])} bytes, budget is {self.budget}"))
            return req
        with self._cv:
            if self._stop:
                raise RuntimeError("server closed")
            self._pending.append(req)
            self._cv.notify_all()
        return req

    def cancel(self, req: Request) -> bool:
        """Remove a request that has not been admitted yet."""
        with self._cv:
            try:
                self._pending.remove(req)
            except ValueError:
                return False
        req.cancelled = True
        req._finish(error=RuntimeError("cancelled"))
        self.stats["cancelled"] += 1
        return True

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        self._exec.close(self._slots)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def in_use(self) -> int:
        return self._in_use

    # ------------------------------------------------------------ scheduling
    @staticmethod
    def _key(r: Request):
        # exact searches fuse only with exact searches: they run another
        # kernel, and a thresholded request would wait for a batch that costs
        # far more than its own
        return (r.kind, r.set_name, r.dtype.str, r.trim, r.kind == "cdist" and r.k < 0)

    def _footprint(self, reqs) -> int:
        out = sum(r.nbytes_out for r in reqs)
        if reqs[0].kind == "pairwise":
            n = sum(r.pairs for r in reqs)
            inp = sum(_span(r.payload[0]) + _span(r.payload[1]) + 16 * (r.pairs + 1) for r in reqs)
            mx = max(max(_longest(r.payload[0]), _longest(r.payload[1])) for r in reqs)
            return out + inp + 4 * n + self._exec.pairs_workspace(n, mx)
        nq = sum(len(r.payload[1]) - 1 for r in reqs)
        inp = sum(_span(r.payload) + 8 * len(r.payload[1]) for r in reqs)
        mq = max(_longest(r.payload) for r in reqs)
        # a length-sorted set computes into a second matrix before the gather
        sorted_set = getattr(self.sets.get(reqs[0].set_name), "col_order", None) is not None
        return (2 * out if sorted_set else out) + inp + 4 * nq + self._exec.cdist_workspace(nq, mq)

    def _take_batch(self):
        """FIFO head and every later pending request it can fuse with."""
        head = self._pending[0]
        key = self._key(head)
        batch, pairs = [], 0
        for r in list(self._pending):
            if self._key(r) != key:
                continue
            if batch and pairs + r.pairs > self.max_batch_pairs:
                break
            batch.append(r)
            pairs += r.pairs
        fp = self._footprint(batch)
        while len(batch) > 1 and self._in_use + fp > self.budget:
            batch.pop()
            fp = self._footprint(batch)
        if self._in_use + fp > self.budget:
            return None, 0
        for r in batch:
            self._pending.remove(r)
        return batch, fp

    def _schedule(self):
        while True:
            with self._cv:
                while not self._stop and not (self._pending and any(not s.busy for s in self._slots)):
                    self._cv.wait()
                if self._stop and not self._pending:
                    return
                if self._stop:
                    # drain: wait for a slot and run what is left
                    while not any(not s.busy for s in self._slots):
                        self._cv.wait()
                batch, fp = self._take_batch()
                if batch is None:
                    self.stats["waited_for_memory"] += 1
                    self._cv.wait(0.05)
                    continue
                slot = next(s for s in self._slots if not s.busy)
                slot.busy = True
                self._in_use += fp
            t = time.perf_counter()
            for r in batch:
                r.t_admit = t
                r.batch_size = len(batch)
            self.stats["batches"] += 1
            self.stats["admitted"] += len(batch)
            threading.Thread(target=self._run, args=(slot, batch, fp), daemon=True).start()

    def _run(self, slot, batch, fp):
        keep_bytes = 0
        try:
            results, out_buf = self._exec.run(self, slot, batch)
            if out_buf is not None:
                # the fused output stays held until the last view is dropped
                keep_bytes = sum(r.nbytes_out for r in batch)
                weakref.finalize(out_buf, self._release, keep_bytes)
            for r, res in zip(batch, results):
                r._finish(result=res)
        except Exception as e:                       # noqa: BLE001 -- handed to the caller
            keep_bytes = 0
            for r in batch:
                r._finish(error=e)
        finally:
            with self._cv:
                self._in_use -= fp - keep_bytes
                slot.busy = False
                self._cv.notify_all()

    def _release(self, nbytes):
        with self._cv:
            self._in_use -= nbytes
            self._cv.notify_all()


class _GpuExecutor:
    """Runs a fused batch on a slot's stream with the batched kernels. Each
    launch takes 1/slots of the GPU's wave slots, so batches of different
    slots run side by side instead of a long one holding every CU while the
    others wait. A search batch whose requests all have a max_dist runs the
    count kernel (the batch's queries as one group against blocks of the
    set, most choices settled by the length and class-count bounds), sized
    so each slot's waves get about two blocks each; exact requests batch
    apart from thresholded ones and run the exact scans (one workgroup per
    query and choice block, the whole Peq table in LDS, 32-row units) over
    the set's stripes in length order."""

    def __init__(self, slots: int = 8, sort_sets: bool = True):
        # sort_sets=False keeps each set in the caller's order (no result
        # gathers): 35-40 % more offline throughput on the serving mix, but
        # at 75 % of that higher rate the pairwise uploads (pageable, over
        # PCIe) saturate and p99 doubles (13-18 -> 24-28 ms)
        self.sort_sets = bool(sort_sets)
        wpc = max(4, 32 // max(1, int(slots)))
        self.wpc = wpc
        self.cdist_params = {"mode": "swar", "waves_per_cu": wpc, "tile": 8192}
        self.pairs_params = {"waves_per_cu": wpc, "block": 256, "bin_width": 16, "chunk": 65536,
                             "trim": True}

    def search_params(self, nq: int, nc: int, kv: np.ndarray) -> dict:
        if len(kv) == 0:
            return self.cdist_params
        if int(kv.min()) < 0:
            # exact scans: one workgroup per (query, choice block) over the
            # set's length-ordered stripes, so a wave's lanes finish together
            return {"mode": "exact"}
        g = min(64, nq)
        waves = self.wpc * 256
        blocks_wanted = max(1, 2 * waves // ((nq + g - 1) // g))
        tile = min(8192, max(256, (nc // blocks_wanted) // 256 * 256))
        return {"mode": "count", "waves_per_cu": self.wpc, "tile": tile, "qgroup": g}

    def new_slot(self, i):
        return _Slot(i)

    def resident(self, packed):
        """A set kept on the device. With sort_sets (the default) the set is
        stored sorted by length (a wave's 64 consecutive choices then have
        similar lengths, so an exact search's lanes finish together) and
        every search's result columns are gathered back into the caller's
        order through col_order (position of each original choice); in the
        serving mix those gathers take 22 % of the GPU time. Without it the
        set stays in the caller's order with both search layouts built once:
        the set-order stripes and profiles (thresholded searches) and stripes
        sorted by length within blocks of hip.SORT_BLOCK strings (exact
        scans, each result written to its choice's own column)."""
        if not self.sort_sets:
            P = hip.Packed(*packed)
            P.col_order = None
            P.stripes()
            P.stripes(order="length")
            call("bl_device_sync")
            return P
        buf, off = packed
        lens = np.diff(off)
        order = np.argsort(lens, kind="stable")
        if len(order) > 1 and bool(np.any(order[1:] < order[:-1])):
            P = hip.Packed(*_gather_strings(buf, off, order))
            inv = np.empty(len(order), np.int64)
            inv[order] = np.arange(len(order), dtype=np.int64)
            P.col_order = DeviceBuffer.from_host(inv)
        else:
            P = hip.Packed(*packed)
            P.col_order = None
        P.stripes()                        # the search layout, built once per set
        call("bl_device_sync")
        return P

    def pairs_workspace(self, n, max_len):
        from ._capi import lib
        return int(lib().bl_pairs_workspace_size(n, max_len))

    def cdist_workspace(self, nq, max_qlen):
        from ._capi import lib
        return int(lib().bl_cdist_workspace_size(nq, min(max_qlen, hip.BATCHED_MAX_QUERY)))

    def close(self, slots):
        for s in slots:
            call("bl_stream_sync", s.stream)
            call("bl_event_destroy", s.event)
            call("bl_stream_destroy", s.stream)

    def run(self, server, slot, batch):
        dt = batch[0].dtype
        if batch[0].kind == "pairwise":
            kv = np.concatenate([np.full(r.pairs, r.k, np.int32) for r in batch])
            A = hip.Packed.from_parts([r.payload[0] for r in batch], slot.stream)
            B = hip.Packed.from_parts([r.payload[1] for r in batch], slot.stream)
            kvec = DeviceBuffer.from_host(kv, slot.stream)
            out = DeviceArray.empty((A.n,), dt)
            hip.pairs_batched_into(A, B, out, -1, batch[0].trim, slot.stream, slot.workspace,
                                   params=self.pairs_params,
                                   kvec=kvec)
            offs = np.cumsum([0] + [r.pairs for r in batch])
            views = [DeviceArray((r.pairs,), dt, out._buf, out._off + int(o) * dt.itemsize)
                     for r, o in zip(batch, offs)]
        else:
            C = server.sets[batch[0].set_name]
            nqs = [len(r.payload[1]) - 1 for r in batch]
            kv = np.concatenate([np.full(q, r.k, np.int32) for r, q in zip(batch, nqs)])
            Q = hip.Packed.from_parts([r.payload for r in batch], slot.stream)
            kvec = DeviceBuffer.from_host(kv, slot.stream)
            out = DeviceArray.empty((Q.n, C.n), dt)
            inv = getattr(C, "col_order", None)
            tmp = DeviceArray.empty((Q.n, C.n), dt) if inv is not None else out
            hip.cdist_batched_into(Q, C, tmp, -1, slot.stream, slot.workspace,
                                   self.search_params(Q.n, C.n, kv), kvec=kvec)
            if inv is not None:
                call("bl_gather_cols", tmp.ptr, inv.ptr, out.ptr, Q.n, C.n, DTYPE_CODE[dt],
                     slot.stream)
            offs = np.cumsum([0] + nqs)
            views = [DeviceArray((q, C.n), dt, out._buf, out._off + int(o) * C.n * dt.itemsize)
                     for q, o in zip(nqs, offs)]
        call("bl_event_record", slot.event, slot.stream)
        call("bl_event_sync", slot.event)
        return views, out._buf


def _resident_bytes(P) -> int:
    """Device bytes a resident set holds (packed buffer, offsets, stripes, and
    the column order of a length-sorted set)."""
    b = getattr(P, "device_bytes", None)
    b = int(b) if b is not None else int(P.nbytes) + 8 * (P.n + 1)
    order = getattr(P, "col_order", None)
    return b + (int(order.nbytes) if order is not None else 0)


def _span(packed) -> int:
    off = packed[1]
    return int(off[-1] - off[0]) if len(off) else 0


def _longest(packed) -> int:
    off = packed[1]
    return int(np.diff(off).max()) if len(off) > 1 else 0


def _gather_strings(buf, off, order, chunk=1 << 16):
    """Packed host set of the strings in `order` (buf, off as packed)."""
    lens = np.diff(off)[order]
    noff = np.zeros(len(order) + 1, np.int64)
    np.cumsum(lens, out=noff[1:])
    nbuf = np.empty(max(int(noff[-1]), 1), np.uint8)
    src = np.asarray(buf, np.uint8)
    for a in range(0, len(order), chunk):
        o = order[a:a + chunk]
        ln = lens[a:a + chunk]
        tot = int(ln.sum())
        if tot == 0:
            continue
        starts = np.repeat(off[o] - noff[a:a + len(o)], ln)
        pos = np.arange(noff[a], noff[a] + tot, dtype=np.int64)
        nbuf[pos] = src[starts + pos]
    return nbuf, noff
