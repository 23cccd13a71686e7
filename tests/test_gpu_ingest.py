"""The staged ingest pipeline (bitlev.ingest): packing on a thread pool into
pinned buffers, copies on their own stream, kernels as copies land. Results
must equal the oracle whatever the chunking, thread count and slot reuse."""
from __future__ import annotations

import numpy as np
import pytest

from bitlev import hip, ingest
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _pairs(seed, n, longest=300):
    rng = np.random.default_rng(seed)
    A, B = [], []
    for i in range(n):
        la = int(rng.integers(0, longest))
        a = bytes(rng.integers(97, 101, la).astype(np.uint8))
        b = bytearray(a)
        for _ in range(int(rng.integers(0, 8))):
            if b:
                b[int(rng.integers(0, len(b)))] = int(rng.integers(97, 101))
        b = bytes(b[: len(b) - int(rng.integers(0, 4))])
        if i % 3 == 0:                       # str items too (code points <= U+00FF)
            A.append(a.decode("latin-1"))
            B.append(b.decode("latin-1"))
        else:
            A.append(a)
            B.append(b)
    return A, B


def _as_bytes(xs):
    return [x.encode("latin-1") if isinstance(x, str) else x for x in xs]


def test_staged_pairs_match_the_oracle_across_chunks_and_threads():
    A, B = _pairs(1, 9000)
    pa, pb = O.pack(_as_bytes(A)), O.pack(_as_bytes(B))
    exact = O.pairs(pa, pb, threads=8)
    for chunk, threads in ((1024, 1), (1024, 3), (4096, 8)):
        for k in (-1, 0, 5):
            for dt in (np.int32, np.uint16, np.uint8):
                got = ingest.pairwise(A, B, k, dt, True, chunk=chunk, threads=threads)
                want = exact if k < 0 else np.minimum(exact, k + 1)
                np.testing.assert_array_equal(got, O.saturate(want, dt),
                                              err_msg=f"chunk={chunk} threads={threads} k={k} {dt}")


def test_staging_buffers_grow_for_later_long_chunks():
    """Early chunks of short strings size the buffers; a later chunk of long
    strings (past the register classes, so the wavefront path too) grows
    them."""
    A, B = _pairs(2, 3000, longest=20)
    rng = np.random.default_rng(3)
    for _ in range(1500):
        a = bytes(rng.integers(97, 101, int(rng.integers(200, 1200))).astype(np.uint8))
        A.append(a)
        B.append(a[5:] + b"xyz")
    A += [b"", b"a"]
    B += [b"abc", b""]
    pa, pb = O.pack(_as_bytes(A)), O.pack(_as_bytes(B))
    got = ingest.pairwise(tuple(A), (x for x in B), -1, np.int32, True, chunk=1024, threads=2)
    np.testing.assert_array_equal(got, O.pairs(pa, pb, threads=8))


def test_public_pairwise_stages_large_calls():
    n = hip.STAGED_PAIRS + 777
    A, B = _pairs(4, n, longest=40)
    pa, pb = O.pack(_as_bytes(A)), O.pack(_as_bytes(B))
    got = hip.pairwise(A, B, max_dist=3, dtype=np.uint8)
    np.testing.assert_array_equal(got, O.saturate(O.pairs(pa, pb, k=3, threads=8), np.uint8))


def test_staged_errors():
    with pytest.raises(ValueError):
        ingest.pairwise([b"a"], [b"a", b"b"])
    with pytest.raises(TypeError):
        ingest.pairwise([b"a"] * 3000, [b"b"] * 2999 + [7], chunk=1024, threads=2)
    with pytest.raises(ValueError):
        ingest.pairwise([b"a"], [b"b"], dtype=np.float32)


def test_cdist_of_long_lists_through_pinned_staging():
    """Choice lists of 32,768 strings or more are packed by the thread pool
    into the reused pinned buffer and uploaded from it; two calls in a row
    (the second with a shorter list) give the oracle's results."""
    rng = np.random.default_rng(8)
    Qs = [bytes(rng.integers(97, 101, m).astype(np.uint8)) for m in (0, 5, 40, 130)]
    for nc in (40000, hip._PARALLEL_PACK):
        C = []
        for i in range(nc):
            s = bytes(rng.integers(97, 101, int(rng.integers(0, 60))).astype(np.uint8))
            C.append(s.decode("latin-1") if i % 2 else s)
        pq, pc = O.pack(Qs), O.pack(_as_bytes(C))
        exact = O.cross(pq, pc, threads=8)
        for k, dt in ((None, np.int32), (3, np.uint8)):
            got = hip.cdist(Qs, C, max_dist=k, dtype=dt)
            want = exact if k is None else np.minimum(exact, k + 1)
            np.testing.assert_array_equal(got, O.saturate(want, dt), err_msg=f"nc={nc} k={k}")
    with pytest.raises(TypeError):
        hip.cdist(Qs, C[:-1] + [None], max_dist=2)
