"""The kernels' per-lane code (csrc/bl_device.h) built for the host and checked
against the golden vectors and the oracle -- no GPU needed. Mirrors the
reference's tests/test_gpu_logic.py:42-122, which does the same for its CUDA
device functions."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from conftest import PKG, cross_of, golden_cross_sets, golden_pair_sets, pairs_of
from oracle import oracle as O

PATH = os.path.join(PKG, "bitlev", "libbl_hostcheck.so")
KS = [-1, 0, 1, 2, 3, 5, 10, 17, 100]
u8p, i64p, i32p = (ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_int64),
                   ctypes.POINTER(ctypes.c_int32))


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(PATH):
        pytest.fail(f"{PATH} not built (make -C editdist-bench_amd)")
    L = ctypes.CDLL(PATH)
    L.blh_pairs.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, i32p, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int]
    L.blh_cross.argtypes = [u8p, i64p, ctypes.c_int64, u8p, i64p, ctypes.c_int64, i32p,
                            ctypes.c_int, ctypes.c_int]
    L.blh_store.argtypes = [ctypes.c_int, ctypes.c_int]
    L.blh_pairs_batched.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, i32p, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int64]
    L.blh_transpose8.argtypes = [ctypes.c_uint64]
    L.blh_transpose8.restype = ctypes.c_uint64
    L.blh_bop3.argtypes = [ctypes.c_uint32] * 3 + [ctypes.c_int]
    L.blh_bop3.restype = ctypes.c_uint32
    L.blh_trim.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_int)]
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def hc_pairs(L, pa, pb, k=-1, trim=1, force_long=0):
    (ba, oa), (bb, ob) = pa, pb
    n = len(oa) - 1
    out = np.empty(n, np.int32)
    L.blh_pairs(_p(ba, u8p), _p(oa, i64p), _p(bb, u8p), _p(ob, i64p), n, _p(out, i32p), trim, k,
                force_long)
    return out


def hc_cross(L, pq, pc, k=-1, force_long=0):
    (bq, oq), (bc, oc) = pq, pc
    nq, nc = len(oq) - 1, len(oc) - 1
    out = np.empty(nq * nc, np.int32)
    L.blh_cross(_p(bq, u8p), _p(oq, i64p), nq, _p(bc, u8p), _p(oc, i64p), nc, _p(out, i32p), k,
                force_long)
    return out.reshape(nq, nc)


@pytest.mark.parametrize("force_long", [0, 1])
def test_pairs_lane_code_matches_golden(hc, golden, force_long):
    for name in golden_pair_sets(golden):
        pa, pb = pairs_of(golden, name)
        for k in KS:
            for trim in (0, 1):
                want = golden[f"pairs/{name}/pairs_k/k{k}/trim{trim}"]
                np.testing.assert_array_equal(hc_pairs(hc, pa, pb, k, trim, force_long), want,
                                              err_msg=f"{name} k={k} trim={trim}")


@pytest.mark.parametrize("force_long", [0, 1])
def test_peq_lane_code_matches_golden(hc, golden, force_long):
    for name in golden_cross_sets(golden):
        pq, pc = cross_of(golden, name)
        for k in KS:
            np.testing.assert_array_equal(hc_cross(hc, pq, pc, k, force_long),
                                          golden[f"cross/{name}/cross_k/k{k}"],
                                          err_msg=f"{name} k={k}")


@pytest.mark.parametrize("bin_width,chunk", [(1, 1024), (16, 2048), (16, 3072), (64, 1 << 20)])
def test_batched_pairs_pipeline_emulated(hc, golden, bin_width, chunk):
    """keys from the offsets -> per-chunk bucket scan -> block-ranked scatter ->
    tiles, as the persistent kernel walks them, with every index and the tile
    word-count invariant checked (small chunks: many chunks, some without long
    pairs or empty)."""
    for name in golden_pair_sets(golden):
        (ba, oa), (bb, ob) = pairs_of(golden, name)
        n = len(oa) - 1
        for k in (-1, 0, 3, 17):
            for trim in (0, 1):
                out = np.full(n, -7, np.int32)
                rc = hc.blh_pairs_batched(_p(ba, u8p), _p(oa, i64p), _p(bb, u8p), _p(ob, i64p), n,
                                          _p(out, i32p), trim, k, bin_width, chunk)
                assert rc == 0, (name, k, trim, rc)
                np.testing.assert_array_equal(out, golden[f"pairs/{name}/pairs_k/k{k}/trim{trim}"],
                                              err_msg=f"{name} k={k} trim={trim}")


@pytest.mark.parametrize("band", [1, 3, 5])
def test_batched_pairs_band_tiles_emulated(hc, golden, band):
    """Band tiles in the emulated pipeline: one window offset per tile from its
    lanes' length differences, the 32-row window when every threshold is <= 14,
    undecided pairs onto the stack and through the whole scan 64 at a time.
    Uneven lengths with edit rates up to 0.45 (so some distances exceed what a
    window reaches), exact and thresholded, plus the golden pair sets."""
    hc.blh_pairs_batched_band.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, i32p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int64)]
    rng = np.random.default_rng(77)
    A, B = [], []
    for _ in range(6000):
        la = int(rng.integers(1, 257))
        a = rng.integers(0, 256 if rng.random() < 0.7 else 4, la).astype(np.uint8)
        rate = float(rng.uniform(0.0, 0.45))
        b = []
        for ch in a.tolist():
            u = rng.random()
            if u >= rate:
                b.append(ch)
            elif u < rate / 3:
                b.append(int(rng.integers(0, 256)))
            elif u < 2 * rate / 3:
                b += [ch, int(rng.integers(0, 256))]
        A.append(bytes(a))
        B.append(bytes(b[:300]))
    sets = [(O.pack(A), O.pack(B), "uneven")]
    sets += [(*pairs_of(golden, name), name) for name in golden_pair_sets(golden)]
    esc = ctypes.c_int64()
    escalated = 0
    for pa, pb, name in sets:
        (ba, oa), (bb, ob) = pa, pb
        n = len(oa) - 1
        for k in (-1, 0, 3, 14, 15, 40):
            for trim in (0, 1):
                want = O.pairs(pa, pb, k=k, trim=bool(trim))
                out = np.full(n, -7, np.int32)
                rc = hc.blh_pairs_batched_band(_p(ba, u8p), _p(oa, i64p), _p(bb, u8p), _p(ob, i64p),
                                               n, _p(out, i32p), trim, k, 8, 1024, band,
                                               ctypes.byref(esc))
                assert rc == 0, (name, k, trim, rc)
                np.testing.assert_array_equal(out, want, err_msg=f"{name} k={k} trim={trim}")
                escalated += esc.value
    assert escalated > 100


def test_wavefront_multi_stripe(hc):
    """Patterns above 4096 bytes take several 64-word stripes and the bottom-row
    scratch between them; compare with the oracle (no word limit)."""
    rng = np.random.default_rng(8)
    A, B = [], []
    for m, extra in [(4095, 3), (4096, 0), (4097, 10), (5000, -7), (8193, 40), (9000, 0)]:
        a = bytes(rng.integers(0, 4, m).astype(np.uint8) + 65)
        b = bytearray(a)
        for _ in range(30):
            b[int(rng.integers(0, len(b)))] = 65 + int(rng.integers(0, 4))
        b = bytes(b[: len(b) + extra]) if extra < 0 else bytes(b) + b"A" * extra
        A.append(a)
        B.append(b)
    A.append(bytes(rng.integers(0, 256, 6000).astype(np.uint8)))
    B.append(bytes(rng.integers(0, 256, 5500).astype(np.uint8)))
    pa, pb = O.pack(A), O.pack(B)
    for k in (-1, 50, 10**6):
        want = O.pairs(pa, pb, k=k)
        np.testing.assert_array_equal(hc_pairs(hc, pa, pb, k, 1, 1), want)
        np.testing.assert_array_equal(hc_pairs(hc, pa, pb, k, 0, 1), want)
    want = O.cross(O.pack(A[:3]), O.pack(B[3:]))
    np.testing.assert_array_equal(hc_cross(hc, O.pack(A[:3]), O.pack(B[3:]), -1, 1), want)


def _near(rng, a, edits, alpha):
    b = bytearray(a)
    for _ in range(edits):
        op = int(rng.integers(0, 3))
        pos = int(rng.integers(0, len(b) + 1))
        ch = int(rng.integers(0, alpha)) + 97
        if op == 0 and pos < len(b):
            b[pos] = ch
        elif op == 1:
            b.insert(pos, ch)
        elif pos < len(b):
            del b[pos]
    return bytes(b)


def test_row_stages_at_boundaries(hc):
    """Thresholded scans run 32 rows, then 64, then all rows; a stage that
    cannot decide hands over when the exit bound's diagonal row leaves it.
    Pattern lengths around every stage edge, texts shorter and longer than
    the pattern by up to the threshold, near duplicates (which survive the
    short stages) and unrelated strings (which exit in them)."""
    rng = np.random.default_rng(33)
    A, B = [], []
    for m in (1, 31, 32, 33, 36, 60, 63, 64, 65, 68, 96, 127, 128, 129, 192, 200, 255, 256):
        for alpha in (2, 4, 26):
            a = bytes(rng.integers(0, alpha, m).astype(np.uint8) + 97)
            for edits in (0, 1, 2, 5, 12, 40):
                A.append(a)
                B.append(_near(rng, a, edits, alpha))
            A.append(a)
            B.append(bytes(rng.integers(0, alpha, max(0, m + int(rng.integers(-8, 9)))).astype(np.uint8) + 97))
    pa, pb = O.pack(A), O.pack(B)
    for k in (0, 1, 2, 3, 4, 8, 16, 28, 31, 40, 64, 100):
        want = O.pairs(pa, pb, k=k)
        for trim in (0, 1):
            np.testing.assert_array_equal(hc_pairs(hc, pa, pb, k, trim), want, err_msg=f"k={k}")
        np.testing.assert_array_equal(np.diag(hc_cross(hc, pa, pb, k)), O.pairs(pa, pb, k=k, trim=False),
                                      err_msg=f"cross k={k}")


def test_32_row_unit_lanes(hc):
    """The batched tiles' lane code in 32-row units (u64 blocks plus a u32 tail
    block for an odd unit count) equals the oracle for every unit count that
    covers the shorter string, exact and thresholded, trimmed or not."""
    hc.blh_pair_w32.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.c_int]
    rng = np.random.default_rng(55)
    cases = []
    for mn in (1, 5, 31, 32, 33, 63, 64, 65, 95, 96, 97, 127, 128, 129, 160, 161, 192, 193, 224,
               225, 255, 256):
        for alpha in (4, 256):
            a = bytes(rng.integers(0, alpha, mn).astype(np.uint8))
            b = bytearray(a)
            for _ in range(int(rng.integers(0, 12))):
                b[int(rng.integers(0, len(b)))] = int(rng.integers(0, alpha))
            b = bytes(b) + bytes(rng.integers(0, alpha, int(rng.integers(0, 9))).astype(np.uint8))
            cases.append((a, b))
            cases.append((b, a))
            cases.append((a, bytes(rng.integers(0, alpha, mn + 3).astype(np.uint8))))
    pa, pb = O.pack([c[0] for c in cases]), O.pack([c[1] for c in cases])
    for k in (-1, 0, 3, 20, 100):
        for trim in (0, 1):
            want = O.pairs(pa, pb, k=k, trim=bool(trim))
            for i, (a, b) in enumerate(cases):
                need = (min(len(a), len(b)) + 31) // 32
                for w in sorted({max(need, 1), 8}):
                    got = hc.blh_pair_w32(a, len(a), b, len(b), trim, k, w)
                    assert got == want[i], (i, len(a), len(b), k, trim, w)


def test_exact_search_unit_lanes(hc):
    """The exact-search kernel's lane (units_lane: the query's Peq table
    word-major as in its LDS copy, the choice read from stripes) equals the
    oracle's cross result for query lengths at and around every 32-row unit
    edge up to 512 (u32 block, u64 blocks, u64 blocks + u32 tail), texts from
    empty to longer than the query, near copies and unrelated strings, exact
    and thresholded. An off-by-one in the unit count, the tail's dword or the
    word-major index fails here."""
    hc.blh_units.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                             ctypes.c_int]
    rng = np.random.default_rng(71)
    Q, C = [], []
    for m in (1, 2, 31, 32, 33, 63, 64, 65, 95, 96, 97, 127, 128, 129, 159, 160, 161, 191, 192,
              193, 255, 256, 257, 300, 383, 384, 385, 449, 480, 511, 512):
        for alpha in (4, 26, 256):
            base = 97 if alpha <= 26 else 0
            q = bytes(rng.integers(0, alpha, m).astype(np.uint8) + base)
            Q.append(q)
            C.append(_near(rng, q, int(rng.integers(0, 6)), min(alpha, 26)))
            C.append(bytes(rng.integers(0, alpha, int(rng.integers(0, m + 20))).astype(np.uint8)
                           + base))
    C.append(b"")
    pq, pc = O.pack(Q), O.pack(C)
    for k in (-1, 0, 2, 7, 40):
        want = O.cross(pq, pc, k=k)
        for i, q in enumerate(Q):
            for j in range(i % 4, len(C), 4):          # a quarter of the grid, every text shape
                got = hc.blh_units(q, len(q), C[j], len(C[j]), k)
                assert got == want[i, j], (len(q), len(C[j]), k, got, want[i, j])


def _dp_cell(a: bytes, b: bytes, i: int, j: int) -> int:
    """Wagner-Fischer D[i][j] of pattern a against text b."""
    prev = list(range(i + 1))
    for jj in range(1, j + 1):
        cur = [jj] + [0] * i
        for ii in range(1, i + 1):
            cur[ii] = min(prev[ii] + 1, cur[ii - 1] + 1, prev[ii - 1] + (a[ii - 1] != b[jj - 1]))
        prev = cur
    return prev[i]


def test_swar_filter_bounds_equal_the_dp(hc):
    """The 4-choice filter's bound for each byte segment is exactly the DP cell
    D[4 + m - n][4] (its row within 1..8), whatever the other three choices
    are: no carry or shift leaks between segments; and the byte-parallel test
    of all four bounds against k = 0..3 agrees with that cell."""
    hc.blh_swar_bounds.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    hc.blh_swar_over.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                 ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    hc.blh_swar_over.restype = ctypes.c_uint32
    rng = np.random.default_rng(44)
    checked = 0
    for trial in range(3000):
        alpha = int(rng.choice([2, 3, 4, 26, 256]))
        m = int(rng.integers(8, 70))
        q = bytes(rng.integers(0, alpha, m).astype(np.uint8))
        ns, texts = [], []
        for s in range(4):
            n = int(np.clip(m + int(rng.integers(-5, 6)), 4, None))
            if rng.random() < 0.5:
                t = bytearray(q[:n] + bytes(max(0, n - m)))
                for _ in range(int(rng.integers(0, 4))):
                    t[int(rng.integers(0, len(t)))] = int(rng.integers(0, alpha))
                t = bytes(t[:n])
            else:
                t = bytes(rng.integers(0, alpha, n).astype(np.uint8))
            ns.append(n)
            texts.append(t)
        packed = b"".join(t[:4] for t in texts)
        nn = (ctypes.c_int * 4)(*ns)
        out = (ctypes.c_int * 4)()
        hc.blh_swar_bounds(q, m, packed, nn, out)
        overs = [hc.blh_swar_over(q, m, packed, nn, k) for k in range(4)]
        for s in range(4):
            i0 = 4 + m - ns[s]
            if 1 <= i0 <= 8:
                cell = _dp_cell(q, texts[s], i0, 4)
                assert out[s] == cell, (trial, s, m, ns[s])
                for k in range(4):        # the SWAR test of the same bound against k
                    assert bool((overs[k] >> (8 * s + 7)) & 1) == (cell > k), (trial, s, k)
                checked += 1
            else:
                assert out[s] == -1
    assert checked > 5000


def _band_cases(rng, count):
    """Pairs whose optimal alignments run along the band's edges: edits
    bunched at one end (insertions at the front push the path off the main
    diagonal before it returns), spread evenly, or none, plus unrelated strings."""
    cases = []
    for _ in range(count):
        alpha = int(rng.choice([2, 4, 26, 256]))
        m = int(rng.integers(1, 300))
        a = bytes(rng.integers(0, alpha, m).astype(np.uint8))
        kind = int(rng.integers(0, 4))
        if kind == 0:
            g = int(rng.integers(0, 40))
            b = bytes(rng.integers(0, alpha, g).astype(np.uint8)) + a
            if rng.random() < 0.5:
                b = b[: max(0, len(b) - int(rng.integers(0, 40)))]
        elif kind == 1:
            b = bytearray(a)
            for _ in range(int(rng.integers(0, max(1, m // 3)))):
                op, pos = int(rng.integers(0, 3)), int(rng.integers(0, len(b) + 1))
                ch = int(rng.integers(0, alpha))
                if op == 0 and pos < len(b):
                    b[pos] = ch
                elif op == 1:
                    b.insert(pos, ch)
                elif pos < len(b):
                    del b[pos]
            b = bytes(b)
        elif kind == 2:
            b = a
        else:
            b = bytes(rng.integers(0, alpha, max(0, m + int(rng.integers(-20, 21)))).astype(np.uint8))
        cases.append((a, b) if rng.random() < 0.5 else (b, a))
    return cases


def test_band_lane(hc):
    """The band lane (a 32- or 64-row window sliding down the diagonals
    [d0, d0 + win - 2], planes built as it goes) returns the exact result
    whenever the window reaches d (or decides the threshold), -1 -- hand over
    to the whole scan -- exactly when it does not, and never a wrong value,
    at offsets that suit the pair, offsets chosen for other lanes of a tile,
    and offsets that barely reach; thresholded scans stop early with k + 1."""
    hc.blh_band.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.POINTER(ctypes.c_int)]
    hc.blh_band_offset.argtypes = [ctypes.c_int] * 3
    rng = np.random.default_rng(66)
    cases = _band_cases(rng, 700)
    pa, pb = O.pack([c[0] for c in cases]), O.pack([c[1] for c in cases])
    exact = O.pairs(pa, pb, k=-1, trim=False)
    reach = ctypes.c_int()
    decided = handed = 0
    for k in (-1, 0, 2, 8, 30, 62, 100):
        for trim in (0, 1):
            want = O.pairs(pa, pb, k=k, trim=bool(trim))
            for i, (a, b) in enumerate(cases):
                dl = -abs(len(a) - len(b))
                for win in (32, 64):
                    best = hc.blh_band_offset(dl, dl, win)
                    offs = {best, best - 2, best + 2, hc.blh_band_offset(dl - 9, dl, win),
                            hc.blh_band_offset(dl, 0, win), dl, 2 - win}
                    for d0 in offs:
                        got = hc.blh_band(a, len(a), b, len(b), trim, k, d0, win, ctypes.byref(reach))
                        r, d = reach.value, int(exact[i])
                        if r == -2:
                            assert got == want[i], (i, k, trim)
                            continue
                        undecidable = r < 0 or (d > r and (k < 0 or k > r))
                        assert (got == -1) == undecidable, (i, len(a), len(b), k, trim, win, d0, r, d, got)
                        if got != -1:
                            assert got == want[i], (i, len(a), len(b), k, trim, win, d0, r, d, got)
                            decided += 1
                        else:
                            handed += 1
    assert decided > 20000 and handed > 2000, (decided, handed)


def test_random_against_oracle(hc):
    rng = np.random.default_rng(21)
    A, B = [], []
    for _ in range(3000):
        alpha = int(rng.choice([2, 4, 26, 256]))
        A.append(bytes(rng.integers(0, alpha, int(rng.integers(0, 300))).astype(np.uint8)))
        B.append(bytes(rng.integers(0, alpha, int(rng.integers(0, 300))).astype(np.uint8)))
    pa, pb = O.pack(A), O.pack(B)
    for k in (-1, 0, 4, 40):
        np.testing.assert_array_equal(hc_pairs(hc, pa, pb, k), O.pairs(pa, pb, k=k))


def test_store_saturates(hc, golden):
    ins = golden["store8/in"]
    assert [hc.blh_store(int(d), 2) for d in ins] == golden["store8/out"].tolist()
    assert [hc.blh_store(d, 1) for d in (0, 65534, 65535, 65536, 10**9)] == [0, 65534, 65535, 65535, 65535]
    assert [hc.blh_store(d, 0) for d in (0, 70000)] == [0, 70000]


def test_transpose8_and_bitop3(hc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        x = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        y = hc.blh_transpose8(x)
        for r in range(8):
            for c in range(8):
                assert (y >> (8 * c + r)) & 1 == (x >> (8 * r + c)) & 1
    for _ in range(200):
        a, b, c = (int(v) for v in rng.integers(0, 2**32, 3))
        M = 0xFFFFFFFF
        assert hc.blh_bop3(a, b, c, 0) == ((a ^ b) | c) & M
        assert hc.blh_bop3(a, b, c, 1) == (a | ~(b | c)) & M
        assert hc.blh_bop3(a, b, c, 2) == (a & ~(b ^ c)) & M


@pytest.mark.parametrize("a,b", [(b"", b""), (b"abc", b"abc"), (b"abcXdef", b"abcYdef"),
                                 (b"aaaa", b"aa"), (b"xabcab", b"abab"), (b"0123456789", b"01234"),
                                 (b"prefix-middle-suffix", b"prefix-MID-suffix"),
                                 (b"ab", b"abab"), (b"abab", b"ab")])
def test_trim_matches_definition(hc, a, b):
    res = (ctypes.c_int * 4)()
    hc.blh_trim(a, len(a), b, len(b), res)
    n = min(len(a), len(b))
    p = 0
    while p < n and a[p] == b[p]:
        p += 1
    s = 0
    while s < n - p and a[len(a) - 1 - s] == b[len(b) - 1 - s]:
        s += 1
    assert list(res) == [p, len(a) - p - s, p, len(b) - p - s]


def test_escalation_list_batches(hc):
    """cdist_list_kernel's batches on the host: near matches and unrelated
    pairs of queries up to 256 bytes against choices from stripes, batches of
    1-64 with mixed length differences, thresholds 0-40; the band (32 rows
    when it reaches every lane, else 64 where it reaches) and the whole-pattern
    fallback give exactly min(d, k + 1) of the oracle, and the band decides
    most near matches."""
    hc.blh_list_batch.argtypes = [u8p, i64p, i32p, u8p, i64p, ctypes.c_int64, i32p, ctypes.c_int,
                                  ctypes.c_int, i32p, ctypes.POINTER(ctypes.c_int)]
    rng = np.random.default_rng(91)
    Qs = [bytes(rng.integers(0, 4, int(rng.integers(1, 257))).astype(np.uint8) + 65) for _ in range(40)]
    Cs = []
    for i in range(600):
        q = Qs[i % len(Qs)]
        if i % 3:
            b = bytearray(q)                      # edits within the queries' alphabet
            for _ in range(int(rng.integers(0, 12))):
                op, pos, ch = int(rng.integers(0, 3)), int(rng.integers(0, len(b) + 1)), 65 + int(rng.integers(0, 4))
                if op == 0 and pos < len(b):
                    b[pos] = ch
                elif op == 1:
                    b.insert(pos, ch)
                elif pos < len(b):
                    del b[pos]
            Cs.append(bytes(b))
        else:
            Cs.append(bytes(rng.integers(0, 4, int(rng.integers(0, 260))).astype(np.uint8) + 65))
    (bq, oq), (bc, oc) = O.pack(Qs), O.pack(Cs)
    banded_total = 0
    for k in (0, 1, 3, 8, 14, 15, 30, 40):
        want = O.cross((bq, oq), (bc, oc), k=k)
        for trial in range(12):
            n = int(rng.integers(1, 65))
            qi = rng.integers(0, len(Qs), n).astype(np.int32)
            ci = np.array([(int(x) * 3 + int(rng.integers(1, 3))) % len(Cs) if trial % 2 else
                           int(rng.integers(0, len(Cs))) for x in qi], np.int32)
            out = np.full(n, -7, np.int32)
            banded = ctypes.c_int()
            assert hc.blh_list_batch(_p(bq, u8p), _p(oq, i64p), _p(qi, i32p), _p(bc, u8p), _p(oc, i64p),
                                     len(Cs), _p(ci, i32p), n, k, _p(out, i32p), ctypes.byref(banded)) == 0
            np.testing.assert_array_equal(out, want[qi, ci], err_msg=f"k={k} trial={trial}")
            banded_total += banded.value
    assert banded_total > 500


def test_class_profiles_and_count_bound(hc):
    """The search kernels' class-count bound: profiles are the saturated counts
    of b & 15; whenever the bound puts a pair above k its exact distance (the
    oracle) is above k; the bound is reached on pairs that make it tight (so
    an off-by-one in the test or the gap term fails here); above 255 bytes the
    counts saturate, which only shrinks L1, and the gap term still holds."""
    hc.blh_class_profile.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
    hc.blh_count_over.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.c_int]
    rng = np.random.default_rng(77)
    for n in (0, 1, 15, 255, 256, 600):
        s = bytes(rng.integers(0, 256, n).astype(np.uint8)) if n < 600 else b"\x13" * 600
        out = (ctypes.c_uint32 * 4)()
        hc.blh_class_profile(s, n, out)
        want = np.minimum(np.bincount(np.frombuffer(s, np.uint8) & 15, minlength=16), 255)
        got = np.frombuffer(bytes(out), np.uint8)
        assert got.tolist() == want.tolist(), n

    def over(a, b, k):
        return bool(hc.blh_count_over(a, len(a), b, len(b), k))

    # soundness against the exact distances
    A, B = [], []
    for _ in range(4000):
        alpha = int(rng.choice([2, 4, 16, 26, 256]))
        m = int(rng.integers(0, 300))
        a = bytes(rng.integers(0, alpha, m).astype(np.uint8))
        if rng.random() < 0.5:
            b = bytearray(a)
            for _ in range(int(rng.integers(0, 12))):
                op = rng.integers(0, 3)
                pos = int(rng.integers(0, len(b) + 1))
                ch = int(rng.integers(0, alpha))
                if op == 0 and pos < len(b):
                    b[pos] = ch
                elif op == 1:
                    b.insert(pos, ch)
                elif pos < len(b):
                    del b[pos]
            b = bytes(b)
        else:
            if rng.random() < 0.3:                # an anagram: same profile, any distance
                b = np.frombuffer(a, np.uint8).copy()
                rng.shuffle(b)
                b = b.tobytes()
            else:
                b = bytes(rng.integers(0, alpha, int(rng.integers(0, 300))).astype(np.uint8))
        A.append(a)
        B.append(b)
    d = O.pairs(O.pack(A), O.pack(B), threads=8)
    pruned = 0
    for a, b, dist in zip(A, B, d):
        for k in (0, 1, 2, 3, 5, 8, 13, 30):
            if over(a, b, k):
                assert dist > k, (a, b, k, dist)
                pruned += 1
    assert pruned > 4000
    # tight cases: d equals the bound, so k = d - 1 is pruned and k = d is not
    tight = [(b"a" * 10, b"b" * 10), (b"acgt" * 8, b"acgt" * 8 + b"xyz"), (b"", b"q" * 7),
             (b"abc", b"abcabc"), (b"\x01\x11\x21", b"\x02\x12\x22"), (b"a" * 255, b"b" * 255)]
    for a, b in tight:
        dist = int(O.pairs(O.pack([a]), O.pack([b]))[0])
        assert over(a, b, dist - 1) and not over(a, b, dist), (a, b, dist)
    # above 255 bytes: saturated counts
    a, b = b"a" * 300, b"b" * 300            # L1' = 510: bound 255 <= d = 300
    assert over(a, b, 254) and not over(a, b, 255)
    a, b = b"a" * 300, b"a" * 290            # counts saturate alike: L1' = 0, gap 10 -> bound 5 (d = 10)
    assert over(a, b, 4) and not over(a, b, 5)
    # same counts in every class, far apart: the bound cannot prune (it is a bound, not a filter)
    assert not over(b"ab" * 20, b"ba" * 20, 0)


def test_two_gram_profiles_and_bound(hc):
    """The 2-gram bound: profiles are the saturated counts of the 16 classes
    4 c2(x) + c2(y), c2(b) = (b >> 1) & 3, over the string's 2-grams; every
    pair it prunes has an oracle distance above k (random, edited, shuffled
    and DNA pairs); it is reached on pairs that make it tight (one
    substitution in the middle of a string moves two 2-grams out and two in;
    a run of one letter against a run of another), so a dropped gap term, a
    factor of 2 for 4 or an off-by-one fails; past 256 bytes (saturated
    counts) the gap term still holds."""
    hc.blh_class2_profile.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
    hc.blh_count2_over.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.c_int]
    rng = np.random.default_rng(78)
    for n in (0, 1, 2, 17, 256, 600):
        s = bytes(rng.integers(0, 256, n).astype(np.uint8)) if n < 600 else b"A" * 600
        out = (ctypes.c_uint32 * 4)()
        hc.blh_class2_profile(s, n, out)
        a = np.frombuffer(s, np.uint8).astype(np.int64)
        cls = 4 * ((a[:-1] >> 1) & 3) + ((a[1:] >> 1) & 3) if n > 1 else np.zeros(0, np.int64)
        want = np.minimum(np.bincount(cls, minlength=16), 255)
        assert np.frombuffer(bytes(out), np.uint8).tolist() == want.tolist(), n

    def over(a, b, k):
        return bool(hc.blh_count2_over(a, len(a), b, len(b), k))

    A, B = [], []
    for _ in range(4000):
        alpha = rng.choice([b"ACGT", b"ab", b"abcdefghijklmnopqrstuvwxyz", bytes(range(256))])
        m = int(rng.integers(0, 300))
        a = bytes(rng.choice(np.frombuffer(alpha, np.uint8), m)) if m else b""
        r = rng.random()
        if r < 0.4:
            b = bytearray(a)
            for _ in range(int(rng.integers(0, 14))):
                op = rng.integers(0, 3)
                pos = int(rng.integers(0, len(b) + 1))
                ch = int(rng.choice(np.frombuffer(alpha, np.uint8)))
                if op == 0 and pos < len(b):
                    b[pos] = ch
                elif op == 1:
                    b.insert(pos, ch)
                elif pos < len(b):
                    del b[pos]
            b = bytes(b)
        elif r < 0.6:
            x = np.frombuffer(a, np.uint8).copy()
            rng.shuffle(x)
            b = x.tobytes()
        else:
            b = bytes(rng.choice(np.frombuffer(alpha, np.uint8), int(rng.integers(0, 300))))
        A.append(a)
        B.append(b)
    d = O.pairs(O.pack(A), O.pack(B), threads=8)
    pruned = 0
    for a, b, dist in zip(A, B, d):
        for k in (0, 1, 2, 3, 5, 8, 13, 30):
            if over(a, b, k):
                assert dist > k, (a, b, k, dist)
                pruned += 1
    assert pruned > 3000
    # tight: d(a, b) = 1 by one substitution in the middle; the 2-gram L1 is 4
    a = b"ACGTACGTAC"
    b = b"ACGTTCGTAC"
    assert int(O.pairs(O.pack([a]), O.pack([b]))[0]) == 1
    prof_a = (ctypes.c_uint32 * 4)(); prof_b = (ctypes.c_uint32 * 4)()
    hc.blh_class2_profile(a, len(a), prof_a); hc.blh_class2_profile(b, len(b), prof_b)
    l1 = int(np.abs(np.frombuffer(bytes(prof_a), np.uint8).astype(int)
                    - np.frombuffer(bytes(prof_b), np.uint8).astype(int)).sum())
    assert l1 == 4 and over(a, b, 0) and not over(a, b, 1)
    # runs of two letters: d = 10; 2-gram L1 = 18 (9 and 9), gap 0 -> bound 4.5
    a, b = b"A" * 10, b"C" * 10
    assert over(a, b, 4) and not over(a, b, 5)
    # lengths differ: a = x, b = x + 6 more; 2-gram L1 >= 6, gap 6 -> bound 3
    a = b"ACGTACGTACGT"
    b = a + b"AAAAAA"
    assert int(O.pairs(O.pack([a]), O.pack([b]))[0]) == 6
    assert not over(a, b, 6)
    # the gap term decides: 9 vs 29 AA 2-grams, L1 = 20, gap 20 -> bound 10 (d = 20)
    a, b = b"A" * 10, b"A" * 30
    assert over(a, b, 9) and not over(a, b, 10)
    # past 256 bytes both counts saturate (L1' = 0); the gap of 20 2-grams gives 5 (d = 20)
    a, b = b"A" * 300, b"A" * 280
    assert over(a, b, 4) and not over(a, b, 5)


def test_three_gram_profiles_and_bound(hc):
    """The 3-gram bound: profiles are the saturated counts of the 64 classes
    16 c2(x) + 4 c2(y) + c2(z) over the string's 3-grams; every pair it prunes
    has an oracle distance above k (random, edited, shuffled and DNA pairs);
    it is reached on pairs that make it tight (one substitution in the middle
    moves three 3-grams out and three in: L1 = 6, bound 1 = d), a run against
    a run, and a length gap that decides, so a factor 3 or 4 for 6, a dropped
    gap term or an off-by-one fails; past 257 bytes (saturated counts) the
    gap term still holds."""
    hc.blh_class3_profile.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
    hc.blh_count3_over.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                   ctypes.c_int]
    rng = np.random.default_rng(79)
    for n in (0, 1, 2, 3, 17, 257, 258, 900):
        s = bytes(rng.integers(0, 256, n).astype(np.uint8)) if n < 900 else b"A" * 900
        out = (ctypes.c_uint32 * 16)()
        hc.blh_class3_profile(s, n, out)
        a = np.frombuffer(s, np.uint8).astype(np.int64)
        c = (a >> 1) & 3
        cls = 16 * c[:-2] + 4 * c[1:-1] + c[2:] if n > 2 else np.zeros(0, np.int64)
        want = np.minimum(np.bincount(cls, minlength=64), 255)
        assert np.frombuffer(bytes(out), np.uint8).tolist() == want.tolist(), n

    def over(a, b, k):
        return bool(hc.blh_count3_over(a, len(a), b, len(b), k))

    A, B = [], []
    for _ in range(4000):
        alpha = rng.choice([b"ACGT", b"ab", b"abcdefghijklmnopqrstuvwxyz", bytes(range(256))])
        m = int(rng.integers(0, 300))
        a = bytes(rng.choice(np.frombuffer(alpha, np.uint8), m)) if m else b""
        r = rng.random()
        if r < 0.4:
            b = bytearray(a)
            for _ in range(int(rng.integers(0, 14))):
                op = rng.integers(0, 3)
                pos = int(rng.integers(0, len(b) + 1))
                ch = int(rng.choice(np.frombuffer(alpha, np.uint8)))
                if op == 0 and pos < len(b):
                    b[pos] = ch
                elif op == 1:
                    b.insert(pos, ch)
                elif pos < len(b):
                    del b[pos]
            b = bytes(b)
        elif r < 0.6:
            x = np.frombuffer(a, np.uint8).copy()
            rng.shuffle(x)
            b = x.tobytes()
        else:
            b = bytes(rng.choice(np.frombuffer(alpha, np.uint8), int(rng.integers(0, 300))))
        A.append(a)
        B.append(b)
    d = O.pairs(O.pack(A), O.pack(B), threads=8)
    pruned = 0
    for a, b, dist in zip(A, B, d):
        for k in (0, 1, 2, 3, 5, 8, 13, 30):
            if over(a, b, k):
                assert dist > k, (a, b, k, dist)
                pruned += 1
    assert pruned > 3000
    # tight: one substitution in the middle (GTA TAC ACG out, GTT TTC TCG in)
    a, b = b"ACGTACGTAC", b"ACGTTCGTAC"
    assert int(O.pairs(O.pack([a]), O.pack([b]))[0]) == 1
    pa, pb = (ctypes.c_uint32 * 16)(), (ctypes.c_uint32 * 16)()
    hc.blh_class3_profile(a, len(a), pa); hc.blh_class3_profile(b, len(b), pb)
    l1 = int(np.abs(np.frombuffer(bytes(pa), np.uint8).astype(int)
                    - np.frombuffer(bytes(pb), np.uint8).astype(int)).sum())
    assert l1 == 6 and over(a, b, 0) and not over(a, b, 1)
    # runs: d = 10; 3-gram L1 = 16, gap 0 -> bound 16 / 6
    a, b = b"A" * 10, b"C" * 10
    assert over(a, b, 2) and not over(a, b, 3)
    # the gap term decides: 8 vs 28 AAA 3-grams, L1 = 20, gap 20 -> bound 40 / 6 (d = 20)
    a, b = b"A" * 10, b"A" * 30
    assert over(a, b, 6) and not over(a, b, 7)
    # past 257 bytes both counts saturate (L1' = 0); the gap of 20 3-grams gives 20 / 6 (d = 20)
    a, b = b"A" * 300, b"A" * 280
    assert over(a, b, 3) and not over(a, b, 4)


def test_count_l2_bound(hc):
    """The matrix-core filter's bound: d > k whenever the squared L2 distance
    of the two byte-class count vectors exceeds 2 k^2. Every pair it prunes
    (random, edited, anagram pairs over several alphabets, 8 thresholds) has
    an oracle distance above k; it is reached on pairs that make it tight (a
    run of one letter against a run of another: d = 10, squared distance
    200 = 2 * 10^2), so a factor 4 for 2 or >= for > fails; one cross-class
    substitution (squared distance 2) is pruned at k = 0 only; it never prunes
    more than the L1 bound does."""
    hc.blh_count_l2_over.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                     ctypes.c_int]
    hc.blh_count_over.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                  ctypes.c_int]

    def over(a, b, k):
        return bool(hc.blh_count_l2_over(a, len(a), b, len(b), k))

    def over_l1(a, b, k):
        return bool(hc.blh_count_over(a, len(a), b, len(b), k))

    rng = np.random.default_rng(79)
    A, B = [], []
    for _ in range(4000):
        alpha = int(rng.choice([2, 4, 16, 26, 256]))
        m = int(rng.integers(0, 80))
        a = bytes(rng.integers(0, alpha, m).astype(np.uint8) + (97 if alpha == 26 else 0))
        r = rng.random()
        if r < 0.5:
            b = bytearray(a)
            for _ in range(int(rng.integers(0, 10))):
                op = rng.integers(0, 3)
                pos = int(rng.integers(0, len(b) + 1))
                ch = int(rng.integers(0, alpha)) + (97 if alpha == 26 else 0)
                if op == 0 and pos < len(b):
                    b[pos] = ch
                elif op == 1:
                    b.insert(pos, ch)
                elif pos < len(b):
                    del b[pos]
            b = bytes(b)
        elif r < 0.7:
            x = np.frombuffer(a, np.uint8).copy()
            rng.shuffle(x)
            b = x.tobytes()
        else:
            b = bytes(rng.integers(0, alpha, int(rng.integers(0, 80))).astype(np.uint8)
                      + (97 if alpha == 26 else 0))
        A.append(a)
        B.append(b)
    d = O.pairs(O.pack(A), O.pack(B), threads=8)
    pruned = 0
    for a, b, dist in zip(A, B, d):
        for k in (0, 1, 2, 3, 5, 8, 13, 30):
            if over(a, b, k):
                assert dist > k, (a, b, k, dist)
                # the L2 norm is at most the L1 norm: the L1 bound prunes it too
                assert over_l1(a, b, k), (a, b, k)
                pruned += 1
    assert pruned > 3000
    a, b = b"a" * 10, b"b" * 10                  # classes 1 and 2: squared distance 200
    assert int(O.pairs(O.pack([a]), O.pack([b]))[0]) == 10
    assert over(a, b, 9) and not over(a, b, 10)
    a, b = b"abcdefgh", b"abcdxfgh"              # one cross-class substitution: squared distance 2
    assert over(a, b, 0) and not over(a, b, 1)
    a, b = b"abcdefgh", b"abcdefghq"             # one insertion: squared distance 1
    assert over(a, b, 0) and not over(a, b, 1)
    assert not over(b"ab" * 20, b"ba" * 20, 0)   # same counts: the bound says nothing
    assert not over(b"a" * 64, b"b" * 64, 64) and over(b"a" * 64, b"b" * 64, 63)
