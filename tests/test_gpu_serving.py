"""The serving loop on the GPU: many requests of both kinds, mixed sizes,
thresholds and dtypes, submitted from several threads while others run; each
result must equal the same request run alone (and the oracle), and requests
must actually have been fused and overlapped."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from bitlev import hip, serving
from bitlev import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sort_sets", [True, False])
def test_each_result_is_the_same_alone_and_in_the_mix(sort_sets):
    """Both set layouts: stored in length order with the result columns
    gathered back, and kept in the caller's order (exact scans over stripes
    sorted within blocks)."""
    rng = np.random.default_rng(11)
    sets = {f"s{i}": W.generate(5, 20000, seed=100 + i) for i in range(2)}
    srv = serving.Server(slots=4, budget_bytes=2 << 30,
                         executor=serving._GpuExecutor(4, sort_sets=sort_sets))
    for name, packed in sets.items():
        srv.add_set(name, packed=packed)
    specs = []
    for i in range(48):
        k = [None, 0, 1, 2, 3, 4][int(rng.integers(0, 6))]
        dt = [np.uint8, np.uint16, np.int32][int(rng.integers(0, 3))]
        if i % 2:
            n = int(rng.integers(1, 3000))
            pa, pb = W.pairs(6, n, seed=1000 + i)
            specs.append(("pairwise", pa, pb, k, dt))
        else:
            name = f"s{int(rng.integers(0, 2))}"
            buf, off = sets[name]
            pick = rng.integers(0, len(off) - 1, int(rng.integers(1, 20)))
            Q = [bytes(buf[off[j]:off[j + 1]]) for j in pick]
            specs.append(("cdist", O.pack(Q), name, k, dt))
    reqs = [None] * len(specs)

    def submit(lo):
        for i in range(lo, len(specs), 4):
            kind, x, y, k, dt = specs[i]
            reqs[i] = (srv.submit_pairwise(x, y, max_dist=k, dtype=dt) if kind == "pairwise"
                       else srv.submit_cdist(x, y, max_dist=k, dtype=dt))

    th = [threading.Thread(target=submit, args=(t,)) for t in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for (kind, x, y, k, dt), r in zip(specs, reqs):
        got = r.result(timeout=120).get()
        kk = -1 if k is None else k
        if kind == "pairwise":
            alone = hip.pairwise(hip_list(x), hip_list(y), max_dist=k, dtype=dt)
            want = O.saturate(O.pairs(x, y, k=kk, threads=8), dt)
        else:
            alone = hip.cdist(hip_list(x), hip_list(sets[y]), max_dist=k, dtype=dt)
            want = O.saturate(O.cross(x, sets[y], k=kk, threads=8), dt)
        assert got.dtype == np.dtype(dt)
        np.testing.assert_array_equal(got, alone)
        np.testing.assert_array_equal(got, want)
    assert srv.stats["admitted"] == len(specs)
    assert max(r.batch_size for r in reqs) > 1 or srv.stats["batches"] < len(specs)
    lat = sorted(r.latency for r in reqs)
    assert lat[0] > 0
    # several batches ran at once on the device
    spans = sorted((r.t_admit, r.t_done) for r in reqs)
    overlap = any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    assert overlap
    srv.close()


def hip_list(packed):
    buf, off = packed
    return [bytes(buf[off[i]:off[i + 1]]) for i in range(len(off) - 1)]


def test_packed_from_parts_equals_the_concatenated_set():
    """Packed.from_parts (each part's bytes copied to its place on the device,
    no host concatenation) holds exactly the set of the parts in turn: parts
    whose offsets do not start at zero (views into a larger set), empty parts
    and empty strings; a search over it equals the oracle on the joined set."""
    rng = np.random.default_rng(21)
    S = [bytes(rng.integers(97, 101, int(rng.integers(0, 40))).astype(np.uint8)) for _ in range(300)]
    whole = O.pack(S)
    buf, off = whole
    cuts = [0, 7, 7, 120, 299, 300]
    parts = [(buf, off[a:b + 1]) for a, b in zip(cuts[:-1], cuts[1:])]
    stream = hip.Stream()
    P = hip.Packed.from_parts(parts, stream.ptr)
    stream.sync()
    assert P.n == len(S) and P.h_off.tolist() == off.tolist()
    got = np.empty(max(int(off[-1]), 1), np.uint8)
    from bitlev._capi import call
    call("bl_memcpy_dtoh", got.ctypes.data, P.d_buf.ptr, int(off[-1]), None)
    call("bl_stream_sync", None)
    assert bytes(got[: off[-1]]) == bytes(buf[: off[-1]])
    Q = hip.Packed(*O.pack(S[:5]))
    out = hip.DeviceArray.empty((Q.n, P.n), np.int32)
    hip.cdist_batched_into(Q, P, out, 3)
    np.testing.assert_array_equal(out.get(), O.cross(O.pack(S[:5]), whole, k=3))
