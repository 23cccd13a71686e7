"""Stream ordering of the host-side uploads, checked on the CPU with the C ABI
replaced by a recorder (no GPU needed).

The package's streams are non-blocking, so they are not ordered with the null
stream: an index list copied in on the null stream could still be in flight
when a kernel on the caller's stream reads it (a plan built without its
stream did exactly that under the benchmark). A plan therefore uploads on the
stream its calls run on, and an upload given no stream has finished before it
returns.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import conftest  # noqa: F401  (import paths)
from bitlev import device, hip

STREAM = 0x5000


@pytest.fixture
def recorder(monkeypatch):
    log = []
    nxt = [1 << 20]

    def fake(name, *args):
        log.append((name, args))
        if name == "bl_malloc":
            args[0].value = nxt[0]
            nxt[0] += args[1]
        return 0

    monkeypatch.setattr(device, "call", fake)
    monkeypatch.setattr(hip, "call", fake)
    monkeypatch.setattr(device, "POOL", device.DevicePool())
    return log


def _packed(strings):
    buf = np.frombuffer(b"".join(strings) or b"\0", np.uint8).copy()
    off = np.zeros(len(strings) + 1, np.int64)
    off[1:] = np.cumsum([len(s) for s in strings])
    return buf, off


def _copies(log):
    return [(i, a) for i, (n, a) in enumerate(log) if n == "bl_memcpy_htod"]


def test_upload_without_a_stream_has_landed_on_return(recorder):
    b = device.DeviceBuffer(64)
    b.upload(np.arange(16, dtype=np.int32))
    (i, args), = _copies(recorder)
    assert args[3] is None
    assert recorder[i + 1] == ("bl_stream_sync", (None,))


def test_upload_on_a_stream_is_ordered_there_and_does_not_block(recorder):
    b = device.DeviceBuffer(64)
    b.upload(np.arange(16, dtype=np.int32), STREAM)
    (i, args), = _copies(recorder)
    assert args[3] == STREAM
    assert all(n != "bl_stream_sync" for n, _ in recorder)


@pytest.mark.parametrize("kind", ["pairs", "cdist"])
def test_plan_index_lists_go_on_the_calls_stream(recorder, kind):
    A = hip.Packed(*_packed([b"a" * n for n in (3, 20, 40, 100, 200, 300, 900)]), stream=STREAM)
    B = hip.Packed(*_packed([b"b" * n for n in (5, 20, 41, 99, 250, 301, 1000)]), stream=STREAM)
    del recorder[:]
    plan = hip.pairs_plan(A, B, STREAM) if kind == "pairs" else hip.cdist_plan(A, B, STREAM)
    copies = _copies(recorder)
    want = sum(b is not None for b in plan.bufs) + (plan.rest_buf is not None)
    assert len(copies) == want > 1
    assert all(a[3] == STREAM for _, a in copies)
    assert all(n != "bl_stream_sync" for n, _ in recorder)


def test_calls_without_a_plan_build_it_on_their_stream(recorder):
    A = hip.Packed(*_packed([b"ab" * 5, b"x" * 30]), stream=STREAM)
    out = device.DeviceArray((2,), np.int32)
    del recorder[:]
    plan = hip.pairs_into(A, A, out, -1, True, STREAM)
    assert plan.bufs[0] is not None or plan.bufs[1] is not None
    assert all(a[3] == STREAM for _, a in _copies(recorder))
    launches = [a for n, a in recorder if n == "bl_pairs_class"]
    assert launches and all(a[-1] == STREAM for a in launches)


def test_recorder_hands_out_distinct_blocks(recorder):
    a, b = device.DeviceBuffer(100), device.DeviceBuffer(100)
    assert a.ptr != b.ptr and isinstance(a.ptr, int)
    assert isinstance(ctypes.c_void_p(a.ptr).value, int)
