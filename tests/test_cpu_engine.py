"""The CPU engine behind the reference's CPU methods (method="scalar" and the
x86 AVX2 "simd" kernels): the reference's own outputs (tests/golden), brute
force against Wagner-Fischer on tiny inputs, every lane class of the SIMD
kernels (16-, 32- and 64-bit lanes, the scalar path past 64 bytes, queries
past the 16-bit lanes' range) against the oracle, the public API with its
methods, dtypes and errors, and that no GPU module calls it. Mirrors the
reference's CPU tests (tests/test_cmyers.py, tests/test_simd.py,
tests/test_lib.py there). No GPU needed."""
from __future__ import annotations

import glob
import itertools
import os
import threading

import numpy as np
import pytest

from conftest import PKG, cross_of, golden_cross_sets, golden_pair_sets, pairs_of
from oracle import oracle as O

import bitlev
from bitlev import _edpy

KS = [-1, 0, 1, 2, 3, 5, 10, 17, 100]


def cpu_pairs(pa, pb, k=-1, trim=True, simd=False):
    out = np.empty(len(pa[1]) - 1, np.int32)
    _edpy.pairs_packed(pa[0], pa[1], pb[0], pb[1], out, trim, k, simd)
    return out


def cpu_cross(pq, pc, k=-1, simd=False):
    out = np.empty((len(pq[1]) - 1, len(pc[1]) - 1), np.int32)
    _edpy.cross_packed(pq[0], pq[1], pc[0], pc[1], out, k, simd)
    return out


@pytest.mark.parametrize("simd", [False, True])
def test_golden_pairs(golden, simd):
    for name in golden_pair_sets(golden):
        pa, pb = pairs_of(golden, name)
        for k in KS:
            for trim in (0, 1):
                want = golden[f"pairs/{name}/pairs_k/k{k}/trim{trim}"]
                np.testing.assert_array_equal(cpu_pairs(pa, pb, k, bool(trim), simd), want,
                                              err_msg=f"{name} k={k} trim={trim}")


@pytest.mark.parametrize("simd", [False, True])
def test_golden_cross(golden, simd):
    for name in golden_cross_sets(golden):
        pq, pc = cross_of(golden, name)
        np.testing.assert_array_equal(cpu_cross(pq, pc, -1, simd), golden[f"cross/{name}/ed_cross"])
        for k in KS:
            np.testing.assert_array_equal(cpu_cross(pq, pc, k, simd), golden[f"cross/{name}/cross_k/k{k}"],
                                          err_msg=f"{name} k={k}")


def test_brute_force_tiny():
    """Every pair of strings over {a, b} up to length 5 against
    Wagner-Fischer, both forms, with and without thresholds."""
    S = [bytes(p) for L in range(6) for p in itertools.product(b"ab", repeat=L)]
    pq = O.pack(S)
    want = np.array([[O.dp(a, b) for b in S] for a in S], np.int32)
    for simd in (False, True):
        np.testing.assert_array_equal(cpu_cross(pq, pq, -1, simd), want)
        for k in (0, 1, 2):
            np.testing.assert_array_equal(cpu_cross(pq, pq, k, simd), np.minimum(want, k + 1))
    A = [a for a in S for b in S]
    B = [b for a in S for b in S]
    pa, pb = O.pack(A), O.pack(B)
    for simd in (False, True):
        np.testing.assert_array_equal(cpu_pairs(pa, pb, -1, True, simd), want.ravel())
        np.testing.assert_array_equal(cpu_pairs(pa, pb, -1, False, simd), want.ravel())


def _mixed(seed, nq, nc, qmax, cmax):
    rng = np.random.default_rng(seed)

    def rs(n):
        a = int(rng.choice([2, 4, 26, 256]))
        return bytes(rng.integers(0, a, n).astype(np.uint8) + (97 if a == 26 else 0))

    Q = [rs(int(rng.integers(0, qmax))) for _ in range(nq)]
    C = [rs(int(rng.integers(0, cmax))) for _ in range(nc)]
    for i in range(0, nc, 5):                     # near copies of queries
        q = bytearray(Q[i % nq])
        if q:
            q[len(q) // 2] ^= 1
        C[i] = bytes(q) + rs(int(rng.integers(0, 3)))
    return Q, C


def test_every_lane_class_against_the_oracle():
    """Choices of 0-130 bytes (16-, 32-, 64-bit lanes and the scalar path, in
    groups of every size), queries of 0-90 bytes, thresholds from 0 to past
    every distance."""
    Q, C = _mixed(3, 37, 611, 90, 131)
    pq, pc = O.pack(Q), O.pack(C)
    for k in (-1, 0, 1, 2, 5, 30, 40000):
        want = O.cross(pq, pc, k=k, threads=8)
        for simd in (False, True):
            np.testing.assert_array_equal(cpu_cross(pq, pc, k, simd), want, err_msg=f"k={k} simd={simd}")


def test_queries_past_the_16_bit_lanes():
    """A 33,000-byte query cannot be scored in 16-bit lanes: it takes the
    scalar path for the short choices, the wider lanes for the rest."""
    rng = np.random.default_rng(4)
    big = bytes(rng.integers(0, 4, 33000).astype(np.uint8))
    Q = [b"ab", big, b"", big[:100]]
    C = [big[i:i + n] for i, n in ((0, 16), (5, 9), (40, 30), (7, 64), (0, 70))] + [b"", b"x"]
    pq, pc = O.pack(Q), O.pack(C)
    for k in (-1, 3, 40000):
        np.testing.assert_array_equal(cpu_cross(pq, pc, k, True), O.cross(pq, pc, k=k))


def test_pairs_every_class_against_the_oracle():
    rng = np.random.default_rng(5)
    A, B = [], []
    for _ in range(6000):
        a = bytes(rng.integers(0, int(rng.choice([2, 4, 256])), int(rng.integers(0, 300))).astype(np.uint8))
        b = bytearray(a[: max(0, len(a) - int(rng.integers(0, 6)))])
        for _ in range(int(rng.integers(0, 8))):
            if b:
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, 256))
        A.append(a)
        B.append(bytes(b) + bytes(rng.integers(0, 256, int(rng.integers(0, 40))).astype(np.uint8)))
    A += [b"x" * 30, b"", b"y" * 64, b"z" * 65, b"q" * 32]
    B += [b"p" * 5000, b"abc", b"y" * 2, b"", b"q" * 33]
    pa, pb = O.pack(A), O.pack(B)
    for trim in (False, True):
        for k in (-1, 0, 3, 50):
            want = O.pairs(pa, pb, k=k, trim=trim, threads=8)
            for simd in (False, True):
                np.testing.assert_array_equal(cpu_pairs(pa, pb, k, trim, simd), want,
                                              err_msg=f"trim={trim} k={k} simd={simd}")


def test_public_cpu_methods():
    """bitlev.pairwise / cdist with the reference's CPU methods
    (tests/test_lib.py:12-69 there): known answers, dtypes, max_dist."""
    for m in ("scalar", "simd"):
        assert bitlev.pairwise(["kitten", "flaw", "", "abc"], ["sitting", "lawn", "abc", "cba"],
                               method=m).tolist() == [3, 2, 3, 2]
        assert bitlev.cdist(["abc"], ["abd", "x"], method=m).tolist() == [[1, 3]]
        r = bitlev.pairwise(["x" * 300], [""], method=m, dtype=np.uint8)
        assert r.dtype == np.uint8 and r[0] == 255
        r = bitlev.pairwise(["x" * 70000], [""], method=m, dtype="uint16")
        assert r.dtype == np.uint16 and r[0] == 65535
        assert bitlev.pairwise(["kitten"], ["sitting"], max_dist=1, method=m).tolist() == [2]
        assert bitlev.cdist(["kitten"], ["sitting", "kitten"], max_dist=1, method=m,
                            dtype=np.uint8).tolist() == [[2, 0]]
        assert bitlev.pairwise([], [], method=m).shape == (0,)
        assert bitlev.cdist([], ["a"], method=m).shape == (0, 1)
        Q, C = _mixed(6, 9, 200, 70, 140)
        for k in (None, 2):
            want = O.cross(O.pack(Q), O.pack(C), k=-1 if k is None else k)
            np.testing.assert_array_equal(bitlev.cdist(Q, C, max_dist=k, method=m), want)


def test_cpu_method_errors():
    with pytest.raises(ValueError):
        bitlev.pairwise(["a"], ["a", "b"], method="scalar")
    with pytest.raises(ValueError):
        bitlev.pairwise(["a"], ["b"], method="neon")
    with pytest.raises(ValueError):
        bitlev.cdist(["a"], ["b"], dtype=np.int64, method="simd")
    with pytest.raises(ValueError):
        bitlev.pairwise(["a"], ["b"], max_dist=-1, method="scalar")
    with pytest.raises(ValueError):
        bitlev.pairwise(["ok"], ["ā"], method="simd")
    with pytest.raises(TypeError):
        bitlev.cdist([3], ["b"], method="scalar")
    with pytest.raises(ValueError):
        _edpy.pairs_simd(["a"], ["b"], np.empty(1, np.int32), True, 3)        # groups 1, 2 or 4
    with pytest.raises(ValueError):
        _edpy.cross(["a", "b"], ["c"], np.empty(1, np.int32))                  # output too small


def test_calls_release_the_gil():
    """Four threads run the SIMD search at once on disjoint rows and get the
    single-call result."""
    Q, C = _mixed(7, 16, 3000, 64, 70)
    pq, pc = O.pack(Q), O.pack(C)
    want = cpu_cross(pq, pc, 2, True)
    got = np.empty_like(want)

    def run(lo, hi):
        sub = (pq[0], pq[1][lo:hi + 1].copy())
        out = np.empty((hi - lo, len(C)), np.int32)
        _edpy.cross_packed(sub[0], sub[1], pc[0], pc[1], out, 2, True)
        got[lo:hi] = out

    th = [threading.Thread(target=run, args=(4 * i, 4 * i + 4)) for i in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    np.testing.assert_array_equal(got, want)


def test_gpu_modules_never_call_the_cpu_engine():
    """The GPU API has no CPU fallback: the engine is reached only through
    the top-level CPU methods."""
    names = ("pairs_simd", "cross_simd", "pairs_packed", "cross_packed", "_edpy.pairs(", "_edpy.cross(")
    for f in glob.glob(os.path.join(PKG, "bitlev", "*.py")):
        if os.path.basename(f) == "__init__.py":
            continue
        text = open(f).read()
        for n in names:
            assert n not in text, (f, n)
