"""Pin the CPU oracle (oracle/myers_oracle.c) before trusting it.

Checks, in order of independence from the code under test:
  * golden vectors written by the reference's own compiled C core
    (tests/golden/make_golden.py) -- every pair/cross set, every k, trim on/off;
  * known answers from the reference's tests (tests/test_cmyers.py:27-31,
    tests/test_lib.py:30-34, tests/test_phase3b.py:58-60, tests/test_lib_gpu.py:30-33);
  * brute force: every pair of strings over {a, b} up to length 6 against
    Wagner-Fischer, and random tiny pairs against the unbounded-integer Myers;
  * invariants: identity, symmetry, triangle inequality, length bounds, trim
    invariance, pairs == cross diagonal;
  * planted mistakes (ORC_MUTANT builds of the oracle): each must fail a check.
"""
from __future__ import annotations

import ctypes
import itertools
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, cross_of, golden_cross_sets, golden_pair_sets, pairs_of, unpack
from oracle import oracle as O

KS = [-1, 0, 1, 2, 3, 5, 10, 17, 100]


def test_golden_pairs_every_k_and_trim(golden):
    for name in golden_pair_sets(golden):
        pa, pb = pairs_of(golden, name)
        for k in KS:
            for trim in (0, 1):
                want = golden[f"pairs/{name}/pairs_k/k{k}/trim{trim}"]
                got = O.pairs(pa, pb, k=k, trim=bool(trim))
                np.testing.assert_array_equal(got, want, err_msg=f"{name} k={k} trim={trim}")
        for trim in (0, 1):
            np.testing.assert_array_equal(O.pairs(pa, pb, trim=bool(trim)),
                                          golden[f"pairs/{name}/ed_pairs_opt/trim{trim}"])


def test_golden_cross_every_k(golden):
    for name in golden_cross_sets(golden):
        pq, pc = cross_of(golden, name)
        np.testing.assert_array_equal(O.cross(pq, pc), golden[f"cross/{name}/ed_cross"])
        for k in KS:
            np.testing.assert_array_equal(O.cross(pq, pc, k=k), golden[f"cross/{name}/cross_k/k{k}"],
                                          err_msg=f"{name} k={k}")


def test_golden_reference_gpu_functions_agree(golden):
    """The reference's GPU device functions (host build) agree with its C core,
    where they apply (shorter string <= 256 for pairs, query <= 512 for Peq)."""
    for name in golden_pair_sets(golden):
        pa, pb = pairs_of(golden, name)
        la, lb = np.diff(pa[1]), np.diff(pb[1])
        m = np.minimum(la, lb)
        for k in (-1, 0, 2, 5, 17):
            exact = golden[f"pairs/{name}/pairs_k/k{k}/trim1"]
            for trim in (0, 1):
                sw = golden[f"pairs/{name}/gpu_swar/k{k}/trim{trim}"]
                np.testing.assert_array_equal(sw[m <= 256], exact[m <= 256])
                assert (sw[m > 256] == -1).all()
            pq = golden[f"pairs/{name}/gpu_peq/k{k}"]
            ok = la <= 512
            np.testing.assert_array_equal(pq[ok], O.pairs(pa, pb, k=k, trim=False)[ok])


def test_known_answers():
    A = [b"", b"", b"abc", b"kitten", b"flaw", b"intention", b"abc", b"caf\xe9", b"\xff\x00"]
    B = [b"", b"abc", b"", b"sitting", b"lawn", b"execution", b"cba", b"cafe", b"\x00\xff"]
    assert O.pairs(O.pack(A), O.pack(B)).tolist() == [0, 3, 3, 3, 2, 5, 2, 1, 2]
    assert O.pairs(O.pack([b"kitten"]), O.pack([b"sitting"]), k=1).tolist() == [2]
    assert O.cross(O.pack([b"abc"]), O.pack([b"abd", b"x"])).tolist() == [[1, 3]]
    assert O.pairs(O.pack([b"a" * 70]), O.pack([b"b" * 70])).tolist() == [70]


def test_saturation_vector(golden):
    d = golden["store8/in"]
    np.testing.assert_array_equal(O.saturate(d, np.uint8), golden["store8/out"])
    assert O.saturate(np.array([0, 7, 254, 255, 256, 9999]), np.uint8).tolist() == [0, 7, 254, 255, 255, 255]
    assert O.saturate(np.array([65534, 65535, 65536, 10**9]), np.uint16).tolist() == [65534, 65535, 65535, 65535]
    assert O.saturate(np.array([5, 10**9]), np.int32).tolist() == [5, 10**9]


def all_strings(alphabet: bytes, maxlen: int):
    for n in range(maxlen + 1):
        for t in itertools.product(alphabet, repeat=n):
            yield bytes(t)


def test_brute_force_binary_up_to_6():
    S = list(all_strings(b"ab", 6))
    A = [a for a in S for _ in S]
    B = [b for _ in S for b in S]
    want = np.array([O.dp(a, b) for a, b in zip(A, B)], np.int32)
    pa, pb = O.pack(A), O.pack(B)
    for trim in (False, True):
        np.testing.assert_array_equal(O.pairs(pa, pb, trim=trim), want)
    np.testing.assert_array_equal(O.dp_pairs(pa, pb), want)
    for k in (0, 1, 2, 4):
        np.testing.assert_array_equal(O.pairs(pa, pb, k=k), np.minimum(want, k + 1))
    np.testing.assert_array_equal(O.cross(O.pack(S), O.pack(S)).ravel(), want)


def test_bigint_myers_at_word_boundaries():
    rng = np.random.default_rng(5)
    A, B = [], []
    for n in [1, 2, 31, 32, 33, 63, 64, 65, 127, 128, 129, 191, 192, 193, 255, 256, 257, 300]:
        for _ in range(6):
            A.append(bytes(rng.integers(0, 4, n).astype(np.uint8) + 65))
            B.append(bytes(rng.integers(0, 4, int(rng.integers(0, n + 40))).astype(np.uint8) + 65))
    want = [O.bigint_myers(a, b) for a, b in zip(A, B)]
    assert O.pairs(O.pack(A), O.pack(B), trim=False).tolist() == want
    assert O.pairs(O.pack(B), O.pack(A), trim=True).tolist() == want
    assert O.dp_pairs(O.pack(A), O.pack(B)).tolist() == want


def test_invariants():
    rng = np.random.default_rng(11)
    n = 400
    X = [bytes(rng.integers(0, 4, int(rng.integers(0, 90))).astype(np.uint8)) for _ in range(3 * n)]
    a, b, c = X[:n], X[n:2 * n], X[2 * n:]
    P = O.pack
    dab = O.pairs(P(a), P(b))
    assert (O.pairs(P(a), P(a)) == 0).all()
    np.testing.assert_array_equal(dab, O.pairs(P(b), P(a)))
    assert (dab <= O.pairs(P(a), P(c)) + O.pairs(P(c), P(b))).all()
    la, lb = np.array([len(x) for x in a]), np.array([len(x) for x in b])
    assert (np.abs(la - lb) <= dab).all() and (dab <= np.maximum(la, lb)).all()
    assert ((dab == 0) == np.array([x == y for x, y in zip(a, b)])).all()
    np.testing.assert_array_equal(dab, O.pairs(P(a), P(b), trim=False))
    np.testing.assert_array_equal(np.diag(O.cross(P(a[:60]), P(b[:60]))), dab[:60])
    # common prefix/suffix do not change the distance
    pre, suf = bytes([1, 2, 3]) * 7, bytes([3, 0]) * 5
    np.testing.assert_array_equal(O.pairs(P([pre + x + suf for x in a]), P([pre + y + suf for y in b])), dab)


def test_threads_do_not_change_results():
    rng = np.random.default_rng(3)
    A = [bytes(rng.integers(0, 256, int(rng.integers(0, 200))).astype(np.uint8)) for _ in range(3000)]
    B = [bytes(rng.integers(0, 256, int(rng.integers(0, 200))).astype(np.uint8)) for _ in range(3000)]
    for k in (-1, 3):
        np.testing.assert_array_equal(O.pairs(O.pack(A), O.pack(B), k=k, threads=4),
                                      O.pairs(O.pack(A), O.pack(B), k=k))
        np.testing.assert_array_equal(O.cross(O.pack(A[:20]), O.pack(B), k=k, threads=4),
                                      O.cross(O.pack(A[:20]), O.pack(B), k=k))


def test_band_columns_follow_exit_rule():
    """cols = first j+1 where score - (n-1-j) > k, 0 for the length gap, n otherwise."""
    q = [b"kitten"]
    c = [b"sitting", b"kitten", b"xxxxxxxxxxxxxxxxxxxxxxxxxx", b"zzzzzz", b""]
    d, cols = O.cross(O.pack(q), O.pack(c), k=1, cols=True)
    assert d.tolist() == [[2, 0, 2, 2, 2]]
    assert cols[0, 2] == 0 and cols[0, 4] == 0          # |n - m| > k: no column
    assert cols[0, 1] == 6                               # exact match runs to the end
    # brute-force the rule with the DP column minimum of the last row
    for j, t in enumerate(c[:2] + c[3:4]):
        m, n = 6, len(t)
        last = [max(m, 0)]
        for jj in range(1, n + 1):
            last.append(O.dp(q[0], t[:jj]))
        exit_at = next((jj + 1 for jj in range(n) if last[jj + 1] - (n - 1 - jj) > 1), n)
        col = cols[0, [0, 1, 3][j]]
        assert col == exit_at


MUTANTS = {1: "dropped term in Xh", 2: "wrong output row", 3: "carries swapped",
           4: "exit rule off by one", 5: "pair gap test off by one", 6: "empty-query k cap",
           7: "uint8 wraps"}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistakes_are_caught(golden, tmp_path, mutant):
    so = tmp_path / f"m{mutant}.so"
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", f"-DORC_MUTANT={mutant}",
                    "-o", str(so), os.path.join(ROOT, "oracle", "myers_oracle.c"), "-lpthread"],
                   check=True)
    saved = O._lib
    try:
        O._lib = None
        real = O.lib()
        L = ctypes.CDLL(str(so))
        for f in ("orc_pairs", "orc_cross", "orc_dp_pairs", "orc_saturate"):
            getattr(L, f).argtypes = getattr(real, f).argtypes
            getattr(L, f).restype = ctypes.c_int
        O._lib = L
        caught = False
        try:
            test_golden_pairs_every_k_and_trim(golden)
            test_golden_cross_every_k(golden)
            test_saturation_vector(golden)
        except AssertionError:
            caught = True
        assert caught, f"mutant {mutant} ({MUTANTS[mutant]}) passed the golden checks"
    finally:
        O._lib = saved
