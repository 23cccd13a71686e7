"""Host logic of the serving loop (bitlev.serving) with an injected executor
that computes on the CPU oracle: fusion, FIFO order within a fusion class,
admission by memory footprint, rejection, cancellation and memory accounting.
No GPU involved; the GPU serving test is tests/test_gpu_serving.py."""
from __future__ import annotations

import threading
import time
import types

import numpy as np
import pytest

from bitlev import serving
from oracle import oracle as O


class FakeExec:
    """Runs a fused batch with the oracle after a short delay (so requests
    pile up and fuse) and records every batch."""

    def __init__(self, delay=0.02):
        self.delay = delay
        self.batches = []
        self.gate = threading.Event()
        self.gate.set()

    def new_slot(self, i):
        return types.SimpleNamespace(idx=i, busy=False)

    def resident(self, packed):
        buf, off = packed
        return types.SimpleNamespace(n=len(off) - 1, nbytes=int(buf.nbytes), packed=packed)

    def pairs_workspace(self, n, max_len):
        return 1024 + 40 * n

    def cdist_workspace(self, nq, max_qlen):
        return 256 + 2048 * nq

    def close(self, slots):
        pass

    def run(self, server, slot, batch):
        self.gate.wait()
        time.sleep(self.delay)
        self.batches.append([r.id for r in batch])
        res = []
        for r in batch:
            if r.kind == "pairwise":
                d = O.pairs(r.payload[0], r.payload[1], k=r.k, trim=r.trim)
            else:
                d = O.cross(r.payload, server.sets[r.set_name].packed, k=r.k)
            res.append(O.saturate(d, r.dtype))
        return res, None


def words(rng, n, lo=1, hi=20):
    return [bytes(rng.integers(97, 100, int(rng.integers(lo, hi))).astype(np.uint8)) for _ in range(n)]


def test_results_fusion_and_accounting():
    rng = np.random.default_rng(0)
    ex = FakeExec()
    srv = serving.Server(slots=2, budget_bytes=1 << 30, executor=ex)
    C = words(rng, 50)
    srv.add_set("s", C)
    base = srv.in_use
    ex.gate.clear()                         # hold the first batch so others queue up
    reqs, want = [], []
    for i in range(30):
        k = [None, 0, 1, 2][i % 4]
        dt = [np.int32, np.uint8][i % 2]
        if i % 3:
            a, b = words(rng, 7), words(rng, 7)
            reqs.append(srv.submit_pairwise(a, b, max_dist=k, dtype=dt))
            want.append(O.saturate(O.pairs(O.pack(a), O.pack(b), k=-1 if k is None else k), dt))
        else:
            q = words(rng, 3)
            reqs.append(srv.submit_cdist(q, "s", max_dist=k, dtype=dt))
            want.append(O.saturate(O.cross(O.pack(q), O.pack(C), k=-1 if k is None else k), dt))
    ex.gate.set()
    for r, w in zip(reqs, want):
        got = r.result(timeout=30)
        assert got.dtype == w.dtype
        np.testing.assert_array_equal(got, w)
        assert r.latency is not None and r.latency >= 0
    assert max(r.batch_size for r in reqs) > 1          # requests were fused
    for b in ex.batches:                                  # fused batches share dtype/kind
        kinds = {(x.kind, x.dtype.str) for x in reqs if x.id in b}
        assert len(kinds) == 1
    srv.close()
    assert srv.in_use == base


def test_rejects_what_can_never_fit_and_waits_for_memory():
    rng = np.random.default_rng(1)
    ex = FakeExec()
    srv = serving.Server(slots=2, budget_bytes=200_000, executor=ex)
    srv.add_set("s", words(rng, 1000))
    big = srv.submit_cdist(words(rng, 300), "s", dtype=np.int32)     # 1.2 MB of output
    with pytest.raises(MemoryError):
        big.result(timeout=5)
    ok = [srv.submit_cdist(words(rng, 20), "s", dtype=np.int32) for _ in range(6)]
    for r in ok:
        assert r.result(timeout=30).shape == (20, 1000)
    assert srv.stats["rejected"] == 1
    srv.close()


def test_cancel_before_admission():
    rng = np.random.default_rng(2)
    ex = FakeExec(delay=0.05)
    srv = serving.Server(slots=1, budget_bytes=1 << 30, executor=ex)
    ex.gate.clear()
    first = srv.submit_pairwise(words(rng, 3), words(rng, 3))
    time.sleep(0.05)                                      # first is admitted and held
    second = srv.submit_pairwise(words(rng, 3), words(rng, 3), dtype=np.uint16)
    assert srv.cancel(second)
    with pytest.raises(RuntimeError):
        second.result(timeout=1)
    ex.gate.set()
    first.result(timeout=10)
    assert not srv.cancel(first)
    srv.close()


def test_argument_errors():
    srv = serving.Server(slots=1, executor=FakeExec())
    with pytest.raises(KeyError):
        srv.submit_cdist(["a"], "missing")
    with pytest.raises(ValueError):
        srv.submit_pairwise(["a"], ["a", "b"])
    with pytest.raises(ValueError):
        srv.submit_pairwise(["a"], ["b"], max_dist=-1)
    srv.add_set("s", ["x"])
    with pytest.raises(ValueError):
        srv.add_set("s", ["y"])
    with pytest.raises(ValueError):
        srv.submit_cdist(["q" * 600], "s")
    srv.close()


def test_length_sorted_resident_layout():
    """A resident set is reordered by length on the host: every string lands
    intact at its sorted position (empty strings included), and the column
    order maps each original choice to that position."""
    from bitlev.serving import _gather_strings
    rng = np.random.default_rng(3)
    S = [bytes(rng.integers(97, 123, int(rng.integers(0, 40))).astype(np.uint8))
         for _ in range(70000)] + [b""] * 5
    off = np.zeros(len(S) + 1, np.int64)
    np.cumsum([len(x) for x in S], out=off[1:])
    buf = np.frombuffer(b"".join(S), np.uint8)
    order = np.argsort(np.diff(off), kind="stable")
    nbuf, noff = _gather_strings(buf, off, order, chunk=4096)
    assert np.all(np.diff(np.diff(noff)) >= 0)
    for j in range(len(S)):
        assert bytes(nbuf[noff[j]:noff[j + 1]]) == S[order[j]], j
    inv = np.empty(len(order), np.int64)
    inv[order] = np.arange(len(order))
    for i in range(0, len(S), 101):
        j = inv[i]
        assert bytes(nbuf[noff[j]:noff[j + 1]]) == S[i]
