"""Batched form on the GPU: every launch shape the autotuner may pick gives
the same results, HIP-graph replays match the eager call on every replay,
and per-item thresholds equal per-call thresholds."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from bitlev import autotune, hip
from bitlev import workloads as W
from bitlev._capi import call
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mixed_pairs(seed, n):
    rng = np.random.default_rng(seed)
    A, B = [], []
    for _ in range(n):
        la = int(rng.integers(0, 300))
        a = rng.integers(0, 4, la).astype(np.uint8) + 65
        b = a.copy()
        for _ in range(int(rng.integers(0, 20))):
            if len(b):
                b[int(rng.integers(0, len(b)))] = 65 + int(rng.integers(0, 4))
        A.append(bytes(a))
        B.append(bytes(b[: len(b) - int(rng.integers(0, 6))]))
    A += [b"x" * 5000, b"y" * 700]
    B += [b"x" * 4990 + b"abc", b"y" * 650]
    return O.pack(A), O.pack(B)


def test_every_pairs_launch_shape_gives_the_same_results():
    pa, pb = _mixed_pairs(1, 6000)
    A, B = hip.Packed(*pa), hip.Packed(*pb)
    want = {k: O.pairs(pa, pb, k=k, threads=8) for k in (-1, 5)}
    space = {"waves_per_cu": [4, 32], "block": [64, 256], "bin_width": [1, 64],
             "chunk": [256, 1 << 20], "band": [0, 1]}
    direct = {"mode": ["direct"], "waves_per_cu": [1, 32], "block": [64, 256]}
    # lanes per pair (a whole wave above this many units) and words per lane
    lanes = {"waves_per_cu": [32], "chunk": [256, 1 << 20], "band": [0, 3],
             "long_units": [1, 4, 7], "unit64": [False, True]}
    for p in list(autotune._grid(space)) + list(autotune._grid(direct)) + list(autotune._grid(lanes)):
        for k in (-1, 5):
            out = hip.DeviceArray.empty((A.n,), np.int32)
            hip.pairs_batched_into(A, B, out, k, True, None, None, p)
            np.testing.assert_array_equal(out.get(), want[k], err_msg=str(p))


@pytest.mark.parametrize("band", [1, 2, 3, 5])
def test_band_tiles_on_uneven_pairs(band):
    """Band tiles on the uneven workload's kind of pairs (lengths 1-256, bytes
    0-255, edit rates 5-30%) and on harder ones (rates up to 45%, so the window
    leaves some pairs to the whole scan), exact, thresholded with the narrow
    window (k <= 14) and the wide one, per-item thresholds, trimmed or not."""
    A, B = W.pairs(3, 20000, seed=5)
    rng = np.random.default_rng(8)
    Ah, Bh = [], []
    for _ in range(4000):
        a = rng.integers(0, 256, int(rng.integers(1, 257))).astype(np.uint8)
        b = a.copy()
        idx = rng.random(len(b)) < rng.uniform(0.3, 0.45)
        b[idx] = rng.integers(0, 256, int(idx.sum())).astype(np.uint8)
        cut = int(rng.integers(0, 30))
        Ah.append(bytes(a))
        Bh.append(bytes(b[cut:]) + bytes(rng.integers(0, 256, int(rng.integers(0, 30))).astype(np.uint8)))
    p = {"waves_per_cu": 16, "block": 256, "bin_width": 8, "chunk": 16384, "band": band,
         "unit64": band == 3, "long_units": 6 if band == 5 else 0}
    for pa, pb in ((A, B), (O.pack(Ah), O.pack(Bh))):
        X, Y = hip.Packed(*pa), hip.Packed(*pb)
        exact = O.pairs(pa, pb, threads=8)
        for k in (-1, 3, 14, 40):
            for trim in (True, False):
                out = hip.DeviceArray.empty((X.n,), np.int32)
                hip.pairs_batched_into(X, Y, out, k, trim, None, None, p)
                want = exact if k < 0 else np.minimum(exact, k + 1)
                np.testing.assert_array_equal(out.get(), want, err_msg=f"k={k} trim={trim}")
        kv = np.random.default_rng(9).choice([-1, 0, 2, 14, 30], size=X.n).astype(np.int32)
        out = hip.DeviceArray.empty((X.n,), np.int32)
        hip.pairs_batched_into(X, Y, out, -1, True, None, None, p, kvec=hip.DeviceBuffer.from_host(kv))
        np.testing.assert_array_equal(out.get(), np.where(kv < 0, exact, np.minimum(exact, kv + 1)))


@pytest.mark.parametrize("gen,k", [(4, 8), (2, 2), (2, 0), (5, -1)])
def test_every_cdist_launch_shape_gives_the_same_results(gen, k):
    pq, pc = W.cdist(gen, 12, 5000, seed=3)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    want = O.cross(pq, pc, k=k, threads=8)
    space = {"mode": ["strided", "queue", "swar", "gap", "count", "exact"], "waves_per_cu": [4, 32],
             "tile": [64, 1000, 8192]}
    shapes = list(autotune._grid(space)) + [{"mode": "count", "tile": 1000, "list": False}]
    for p in shapes:
        out = hip.DeviceArray.empty((Q.n, C.n), np.uint8)
        hip.cdist_batched_into(Q, C, out, k, None, None, p)
        np.testing.assert_array_equal(out.get(), O.saturate(want, np.uint8), err_msg=str(p))


def test_cdist_stage_escalation_paths():
    """Queue mode hands a choice the 32-row stage cannot decide to the wider
    stages: at once when even its first test needs rows past 32 (queries above
    36 bytes with much shorter choices), after some columns otherwise, and in
    drains of 64 when whole tiles escalate (a threshold no prefix can exceed).
    Queries of at most 32 bytes never escalate."""
    rng = np.random.default_rng(12)
    Qs = [bytes(rng.integers(0, 4, m).astype(np.uint8) + 97) for m in (5, 32, 33, 40, 64, 100, 200, 256)]
    Cs = []
    for i in range(3000):
        q = Qs[i % len(Qs)]
        if i % 3 == 0:
            Cs.append(q[: max(0, len(q) - int(rng.integers(0, 60)))])
        elif i % 3 == 1:
            b = bytearray(q)
            for _ in range(int(rng.integers(0, 6))):
                b[int(rng.integers(0, len(b)))] = 97 + int(rng.integers(0, 4))
            Cs.append(bytes(b))
        else:
            Cs.append(bytes(rng.integers(0, 4, int(rng.integers(0, 260))).astype(np.uint8) + 97))
    pq, pc = O.pack(Qs), O.pack(Cs)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    for k in (0, 1, 2, 3, 8, 30, 60, 300):
        want = O.cross(pq, pc, k=k, threads=8)
        for p in ({"mode": "queue", "tile": 512}, {"mode": "queue", "tile": 8192},
                  {"mode": "strided", "tile": 512}, {"mode": "swar", "tile": 320},
                  {"mode": "swar", "tile": 8192}):
            out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
            hip.cdist_batched_into(Q, C, out, k, None, None, p)
            np.testing.assert_array_equal(out.get(), want, err_msg=f"k={k} {p}")


def test_gap_windows_edge_cases():
    """Gap windows: queries no longer than their threshold (an empty choice is
    then in band and its scan returns m), empty choices, lengths around the
    gap, windows cut by the tile end, exact queries (which run another mode),
    every output dtype."""
    rng = np.random.default_rng(31)
    Qs = [b"a", b"ab", b"abc", bytes(rng.integers(97, 101, 40).astype(np.uint8)),
          bytes(rng.integers(97, 101, 200).astype(np.uint8))]
    Cs = []
    for i in range(3001):
        q = Qs[i % len(Qs)]
        r = i % 7
        if r == 0:
            Cs.append(b"")
        elif r < 4:
            b = bytearray(q[: max(0, len(q) - int(rng.integers(0, 4)))])
            for _ in range(int(rng.integers(0, 3))):
                if b:
                    b[int(rng.integers(0, len(b)))] = int(rng.integers(97, 101))
            Cs.append(bytes(b))
        else:
            Cs.append(bytes(rng.integers(97, 101, int(rng.integers(0, 60))).astype(np.uint8)))
    pq, pc = O.pack(Qs), O.pack(Cs)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    for k in (0, 1, 3, 5, 8, 40):
        for tile in (256, 1024):
            for dt in (np.uint8, np.uint16, np.int32):
                out = hip.DeviceArray.empty((Q.n, C.n), dt)
                hip.cdist_batched_into(Q, C, out, k, None, None, {"mode": "gap", "tile": tile})
                np.testing.assert_array_equal(out.get(), O.saturate(np.minimum(exact, k + 1), dt),
                                              err_msg=f"k={k} tile={tile} {dt}")
    kv = np.array([-1, 0, 4, -1, 8], np.int32)
    out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
    hip.cdist_batched_into(Q, C, out, -1, None, None, {"mode": "gap"}, kvec=hip.DeviceBuffer.from_host(kv))
    np.testing.assert_array_equal(out.get(), np.where(kv[:, None] < 0, exact, np.minimum(exact, kv[:, None] + 1)))


def test_class_count_bound_in_searches():
    """The class-count bound the batched searches apply before any scan:
    choices built to sit exactly at the bound (edits between distinct byte
    classes, d = k and k + 1), anagrams of the query (same profile, large
    distance), choices differing from the query only inside one class
    (profile equal, d > 0), queries of 255 and 256 bytes of one byte (the
    count saturates at 256), empty choices; the filter, gap and compacted
    modes, dense and match-list output, with the class and 2-gram profiles,
    with the class profiles only and without profiles."""
    rng = np.random.default_rng(53)
    base = bytes(rng.integers(97, 123, 64).astype(np.uint8))
    Qs = [base, b"\x41" * 255, b"\x41" * 256, base[:20], bytes(rng.integers(0, 256, 200).astype(np.uint8)),
          b""]
    Cs = []
    for i in range(4096):
        q = Qs[i % (len(Qs) - 1)]
        r = i % 9
        b = bytearray(q)
        if r == 0:
            Cs.append(b"")
            continue
        if r == 1:                              # anagram
            a = np.frombuffer(q, np.uint8).copy()
            rng.shuffle(a)
            Cs.append(a.tobytes())
            continue
        if r == 2:                              # same class: b ^ 0x10 keeps b & 15
            for _ in range(int(rng.integers(1, 5))):
                p_ = int(rng.integers(0, len(b)))
                b[p_] ^= 0x10
        elif r in (3, 4, 5):                    # t edits across classes
            for _ in range(int(rng.integers(0, 6))):
                op = int(rng.integers(0, 3))
                p_ = int(rng.integers(0, len(b) + 1))
                ch = int(rng.integers(0, 256))
                if op == 0 and p_ < len(b):
                    b[p_] = ch
                elif op == 1:
                    b.insert(p_, ch)
                elif p_ < len(b):
                    del b[p_]
        elif r == 6:                            # one byte value, around 255/256 long
            b = bytearray(b"\x41" * int(rng.integers(250, 258)))
        elif r == 7 and i % 2:                  # far longer than any query: gap > k up to 300
            b = bytearray(rng.integers(0, 256, int(rng.integers(300, 500))).astype(np.uint8))
        else:
            b = bytearray(rng.integers(0, 256, int(rng.integers(0, 260))).astype(np.uint8))
        Cs.append(bytes(b))
    pq, pc = O.pack(Qs), O.pack(Cs)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    S = C.stripes(grams3=True)
    assert S.desc.counts and S.desc.counts2 and S.desc.counts3
    saved, saved2, saved3 = S.desc.counts, S.desc.counts2, S.desc.counts3
    try:
        for with_profiles, with_2grams, with_3grams in ((True, True, True), (True, True, False),
                                                        (True, False, False), (False, False, False)):
            S.desc.counts = saved if with_profiles else None
            S.desc.counts2 = saved2 if with_2grams else None
            S.desc.counts3 = saved3 if with_3grams else None
            for k in (0, 1, 2, 3, 5, 8, 300):
                want = np.minimum(exact, k + 1)
                for mode, extra in (("swar", {}), ("gap", {}), ("compacted", {}), ("count", {}),
                                    ("count", {"qgroup": 2}), ("count", {"qgroup": 64})):
                    for dt in (np.uint8, np.uint16, np.int32):
                        out = hip.DeviceArray.empty((Q.n, C.n), dt)
                        hip.cdist_batched_into(Q, C, out, k, None, None, {"mode": mode, "tile": 1024, **extra})
                        np.testing.assert_array_equal(out.get(), O.saturate(want, dt),
                                                      err_msg=f"k={k} {mode} {extra} {dt} profiles={with_profiles}")
                for mode in ("swar", "count"):
                    qi, ci, di = hip.cdist_matches_into(Q, C, k, params={"mode": mode})
                    got = sorted(zip(qi.get().tolist(), ci.get().tolist(), di.get().tolist()))
                    wq, wc = np.nonzero(exact <= k)
                    assert got == sorted(zip(wq.tolist(), wc.tolist(), exact[wq, wc].tolist())), \
                        (k, mode, with_profiles)
    finally:
        S.desc.counts, S.desc.counts2, S.desc.counts3 = saved, saved2, saved3


def _want_profile(s, width):
    """Saturated class counts of the width-byte windows of s (1: byte class
    b & 15; 2: 4 c2(x) + c2(y); 3: 16 c2(x) + 4 c2(y) + c2(z), c2(b) = (b >> 1) & 3)."""
    a = np.frombuffer(s, np.uint8).astype(np.int64)
    if width == 1:
        cls, size = a & 15, 16
    else:
        c2 = (a >> 1) & 3
        if len(a) < width:
            return np.zeros(16 if width == 2 else 64, np.uint8)
        if width == 2:
            cls, size = 4 * c2[:-1] + c2[1:], 16
        else:
            cls, size = 16 * c2[:-2] + 4 * c2[1:-1] + c2[2:], 64
    return np.minimum(np.bincount(cls, minlength=size), 255).astype(np.uint8)


@pytest.mark.parametrize("order", ["set", "length"])
def test_stripes_hold_every_string_and_profile(order):
    """Words, lengths and profiles of a built set against a restatement:
    groups whose bytes fit the build's LDS stage and groups that do not (a
    20 kB string), empty strings, a ragged last group, 70,000 bytes of one
    class (a count past 16 bits), and the length-ordered layout."""
    rng = np.random.default_rng(11)
    strs = []
    for i in range(64 * 37 + 5):
        r = i % 9
        if r == 0:
            strs.append(b"")
        elif r in (1, 2, 3):
            strs.append(bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), int(rng.integers(1, 300)))))
        else:
            strs.append(bytes(rng.integers(0, 256, int(rng.integers(1, 260))).astype(np.uint8)))
    strs[64 * 5 + 3] = bytes(rng.integers(0, 256, 20000).astype(np.uint8))
    strs[64 * 9 + 63] = b"A" * 70000
    strs[64 * 9 + 62] = b"AC" * 40000
    C = hip.Packed(*O.pack(strs))
    n = C.n
    if order == "set":
        S = C.stripes(grams3=True)
        slot = np.arange(n)
    else:
        S = C.stripes(order="length")
        assert S.d_order is not None
        slot = hip.DeviceArray((n,), np.int32, S.d_order).get().astype(np.int64)
        assert sorted(slot.tolist()) == list(range(n))
    groups = (n + 63) // 64
    gbase = hip.DeviceArray((groups + 1,), np.int64, S.d_gbase).get()
    words = hip.DeviceArray((int(gbase[-1]),), np.uint32, S.d_words).get()
    lens = hip.DeviceArray((n,), np.int32, S.d_len).get()
    have = {}
    if S.d_counts is not None:
        have[1] = hip.DeviceArray((n, 16), np.uint8, S.d_counts).get()
        have[2] = hip.DeviceArray((n, 16), np.uint8, S.d_counts2).get()
    if S.d_counts3 is not None:
        have[3] = hip.DeviceArray((n, 64), np.uint8, S.d_counts3).get()
    assert (order == "set") == (1 in have and 3 in have)
    for g in range(groups):
        width = int(gbase[g + 1] - gbase[g]) // 64
        members = [strs[slot[c]] for c in range(64 * g, min(64 * g + 64, n))]
        assert width == max(1, max((len(x) + 3) // 4 for x in members)), g
        block = words[gbase[g]: gbase[g + 1]].reshape(width, 64)
        for lane, x in enumerate(members):
            c = 64 * g + lane
            assert lens[c] == len(x), c
            padded = np.zeros(4 * width, np.uint8)
            padded[: len(x)] = np.frombuffer(x, np.uint8)
            np.testing.assert_array_equal(block[:, lane], padded.view("<u4"), err_msg=f"slot {c}")
            for w, prof in have.items():
                np.testing.assert_array_equal(prof[c], _want_profile(x, w), err_msg=f"slot {c} grams {w}")
        for lane in range(len(members), 64):
            assert not block[:, lane].any(), g


@pytest.mark.parametrize("shift", [1, 3, 5, 8, 13, 15])
def test_stripes_build_from_an_unaligned_buffer(shift):
    """The packed bytes may start anywhere (a caller's sub-buffer): the build
    stages each group from the aligned 16-byte block holding its first byte,
    so the words and profiles must not depend on the buffer's alignment."""
    rng = np.random.default_rng(shift)
    strs = [bytes(rng.integers(0, 256, int(rng.integers(0, 120))).astype(np.uint8))
            for _ in range(64 * 5 + 17)]
    buf, off = O.pack(strs)
    n = len(strs)
    d_buf = hip.DeviceBuffer(len(buf) + 32)
    d_buf.upload(np.frombuffer(bytes(buf), np.uint8), None, shift)
    d_off = hip.DeviceBuffer.from_host(np.asarray(off, np.int64))
    groups = (n + 63) // 64
    gbase = np.empty(groups + 1, np.int64)
    total = call("bl_stripes_layout", np.asarray(off, np.int64).ctypes.data, n, gbase.ctypes.data)
    d_gbase = hip.DeviceBuffer.from_host(gbase)
    d_words, d_len = hip.DeviceBuffer(4 * total), hip.DeviceBuffer(4 * n)
    d_c1, d_c2, d_c3 = hip.DeviceBuffer(16 * n), hip.DeviceBuffer(16 * n), hip.DeviceBuffer(64 * n)
    call("bl_stripes_build_order", d_buf.ptr + shift, d_off.ptr, n, None, d_gbase.ptr, d_words.ptr,
         d_len.ptr, d_c1.ptr, d_c2.ptr, d_c3.ptr, None)
    call("bl_stream_sync", None)
    words = hip.DeviceArray((total,), np.uint32, d_words).get()
    prof = {1: hip.DeviceArray((n, 16), np.uint8, d_c1).get(),
            2: hip.DeviceArray((n, 16), np.uint8, d_c2).get(),
            3: hip.DeviceArray((n, 64), np.uint8, d_c3).get()}
    assert (hip.DeviceArray((n,), np.int32, d_len).get() == np.diff(off)).all()
    for c, x in enumerate(strs):
        g, lane = divmod(c, 64)
        width = int(gbase[g + 1] - gbase[g]) // 64
        padded = np.zeros(4 * width, np.uint8)
        padded[: len(x)] = np.frombuffer(x, np.uint8)
        np.testing.assert_array_equal(words[gbase[g] + lane: gbase[g + 1]: 64], padded.view("<u4"),
                                      err_msg=f"string {c}")
        for w, p in prof.items():
            np.testing.assert_array_equal(p[c], _want_profile(x, w), err_msg=f"string {c} grams {w}")


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 16384, 16385, 1_000_003])
def test_device_layout_equals_the_host_layout(n):
    """bl_stripes_layout_device against bl_stripes_layout on the same
    offsets: empty strings, a ragged last group, one block of groups and
    more (the block-total scan), a heavy tail of long strings."""
    rng = np.random.default_rng(n)
    lens = rng.integers(0, 70, n)
    tail = rng.random(n) < 0.01
    lens[tail] = rng.integers(1000, 100000, int(tail.sum()))
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(lens)
    want = np.empty((n + 63) // 64 + 1, np.int64)
    total = call("bl_stripes_layout", off.ctypes.data, n, want.ctypes.data)
    d_off = hip.DeviceBuffer.from_host(off)
    d_gbase = hip.DeviceBuffer(8 * len(want))
    work = hip.DeviceBuffer(call("bl_stripes_layout_workspace_size", n))
    got = call("bl_stripes_layout_device", d_off.ptr, n, d_gbase.ptr, work.ptr, work.nbytes, None)
    assert got == total
    np.testing.assert_array_equal(hip.DeviceArray((len(want),), np.int64, d_gbase).get(), want)
    if n >= 2:
        bad = off.copy()
        bad[n // 2] = bad[n // 2 + 1] + 1          # one string of negative length
        hip.DeviceBuffer.upload(d_off, bad)
        with pytest.raises(ValueError, match="offsets decrease"):
            call("bl_stripes_layout_device", d_off.ptr, n, d_gbase.ptr, work.ptr, work.nbytes, None)
    with pytest.raises(ValueError, match="workspace"):
        call("bl_stripes_layout_device", d_off.ptr, n, d_gbase.ptr, work.ptr, 0, None)


def test_filter_mode_with_every_choice_open():
    """The four-choice filter leaves every choice open when all are within k
    of the query (here: copies with at most one substitution, k = 3): the
    rescan stack fills 64 at a time in every window; mixed with per-query
    thresholds that send some queries through the compacted mode (k > 3) and
    queries too short for the filter."""
    rng = np.random.default_rng(17)
    Qs = [bytes(rng.integers(97, 123, m).astype(np.uint8)) for m in (64, 40, 8, 7, 100)]
    Cs = []
    for i in range(4000):
        b = bytearray(Qs[0])
        if i % 2:
            b[int(rng.integers(0, 64))] = 97 + int(rng.integers(0, 26))
        Cs.append(bytes(b[: 64 - (i % 3)]))
    pq, pc = O.pack(Qs), O.pack(Cs)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    for kq in ([3, 3, 3, 3, 3], [3, 9, 2, 1, 0], [0, 1, 2, 3, -1]):
        kv = np.array(kq, np.int32)
        want = np.where(kv[:, None] < 0, exact, np.minimum(exact, kv[:, None] + 1))
        for p in ({"mode": "swar", "tile": 256}, {"mode": "swar", "tile": 8192},
                  {"mode": "swar", "tile": 1024, "waves_per_cu": 1},
                  {"mode": "count", "tile": 256, "qgroup": 3}, {"mode": "count", "waves_per_cu": 1}):
            out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
            hip.cdist_batched_into(Q, C, out, -1, None, None, p, kvec=hip.DeviceBuffer.from_host(kv))
            np.testing.assert_array_equal(out.get(), want, err_msg=f"{kq} {p}")


@pytest.mark.parametrize("mode", ["swar", "queue", "strided", "gap", "count", "exact"])
def test_match_list_equals_the_dense_result(mode):
    """bl_cdist_matches returns exactly the cells the dense cdist puts at or
    below k (here with per-query thresholds, one of them negative: that query
    adds nothing), whatever the lane mode, and grows a too-small buffer."""
    pq, pc = W.cdist(2, 16, 6000, seed=9)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    kv = np.array([2, 3, 0, 1, 8, -1, 2, 2, 3, 4, 0, 1, 2, 3, 5, 40], np.int32)
    for cap in (None, 1):
        dq, dc, dd = hip.cdist_matches_into(Q, C, -1, params={"mode": mode},
                                            kvec=hip.DeviceBuffer.from_host(kv), capacity=cap)
        got = sorted(zip(dq.get().tolist(), dc.get().tolist(), dd.get().tolist()))
        qi, ci = np.nonzero((exact <= kv[:, None]) & (kv[:, None] >= 0))
        want = sorted(zip(qi.tolist(), ci.tolist(), exact[qi, ci].tolist()))
        assert got == want, (mode, cap, len(got), len(want))
    q, c, d = hip.cdist_matches([b"kitten", b"flaw"], [b"sitting", b"lawn", b"kitten"], 2)
    assert (q.tolist(), c.tolist(), d.tolist()) == ([0, 1], [2, 1], [0, 2])


def _exact_call(Q, C, out, k, stripes, params=None, kvec=None):
    """bl_cdist_tuned straight through the C ABI on the given stripes."""
    mq = min(Q.max_len, hip.BATCHED_MAX_QUERY)
    ws = hip.DeviceBuffer(call("bl_cdist_workspace_size", Q.n, mq))
    call("bl_cdist_tuned", Q.d_buf.ptr, Q.d_off.ptr, Q.n, stripes.ref, mq, out.ptr,
         hip.DTYPE_CODE[out.dtype], int(k), kvec.ptr if kvec else None, ws.ptr, ws.nbytes,
         hip._cdist_params(params), None)
    call("bl_stream_sync", None)


@pytest.mark.parametrize("nc", [1, 63, 65, 3000])
def test_exact_scans_every_unit_count(nc):
    """Exact-scan mode (5): queries of 1-512 bytes (every count of 32-row
    units, u64 blocks with and without a u32 tail) against choices of 0-300
    bytes, on the set's own stripes and on stripes in length order (results in
    the caller's columns), in every dtype, exact and with per-query thresholds
    (some exact); the automatic mode of an exact search is this one, and the
    plain stripes kernel refuses a set in length order."""
    rng = np.random.default_rng(nc)
    ms = [1, 2, 31, 32, 33, 64, 65, 96, 97, 128, 129, 160, 161, 200, 256, 257, 300, 383, 400, 511, 512]
    Qs = [bytes(rng.integers(0, 4, m).astype(np.uint8) + 97) for m in ms] + [b""]
    Cs = []
    for i in range(nc):
        q = Qs[i % (len(Qs) - 1)]
        if i % 4 == 0:
            b = bytearray(q[: 300])
            for _ in range(int(rng.integers(0, 8))):
                b[int(rng.integers(0, len(b)))] = 97 + int(rng.integers(0, 4))
            Cs.append(bytes(b))
        else:
            Cs.append(bytes(rng.integers(0, 4, int(rng.integers(0, 301))).astype(np.uint8) + 97))
    pq, pc = O.pack(Qs), O.pack(Cs)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    kq = rng.choice([-1, 0, 3, 17, 300], size=Q.n).astype(np.int32)
    kv = hip.DeviceBuffer.from_host(kq)
    capped = np.where(kq[:, None] < 0, exact, np.minimum(exact, kq[:, None] + 1))
    for order in ("set", "length"):
        st = C.stripes(None, order)
        for tile in (64, 448, 4096):
            for dt in (np.int32, np.uint16, np.uint8):
                out = hip.DeviceArray.empty((Q.n, C.n), dt)
                _exact_call(Q, C, out, -1, st, {"mode": "exact", "tile": tile})
                np.testing.assert_array_equal(out.get(), O.saturate(exact, dt),
                                              err_msg=f"{order} tile={tile} {dt}")
            out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
            _exact_call(Q, C, out, -1, st, {"mode": "exact", "tile": tile}, kvec=kv)
            np.testing.assert_array_equal(out.get(), capped, err_msg=f"{order} tile={tile} kvec")
        out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
        _exact_call(Q, C, out, -1, st)                  # automatic mode
        np.testing.assert_array_equal(out.get(), exact, err_msg=f"{order} default")
    # the public path: an exact search runs the exact scans on the length-ordered set
    out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
    hip.cdist_batched_into(Q, C, out, -1, None)
    np.testing.assert_array_equal(out.get(), exact)
    got = hip.cdist([bytes(q) for q in Qs], [bytes(c) for c in Cs], dtype=np.uint16)
    np.testing.assert_array_equal(got, O.saturate(exact, np.uint16))
    st = C.stripes(None, "length")
    # a set already in length order (one string here) keeps its own stripes
    assert bool(st.desc.order) == (nc > 1)
    if st.desc.order:
        qidx = hip.DeviceBuffer.from_host(np.arange(2, dtype=np.int32))
        with pytest.raises(ValueError):
            call("bl_cdist_class_stripes", 1, Q.d_buf.ptr, Q.d_off.ptr, qidx.ptr, 2, st.ref, out.ptr,
                 hip.DTYPE_CODE[out.dtype], -1, None)


def test_per_item_thresholds():
    pa, pb = _mixed_pairs(2, 3000)
    A, B = hip.Packed(*pa), hip.Packed(*pb)
    rng = np.random.default_rng(4)
    kv = rng.choice([-1, 0, 1, 3, 10], size=A.n).astype(np.int32)
    exact = O.pairs(pa, pb, threads=8)
    for p in (None, {"mode": "direct"}):
        out = hip.DeviceArray.empty((A.n,), np.int32)
        hip.pairs_batched_into(A, B, out, -1, True, None, params=p,
                               kvec=hip.DeviceBuffer.from_host(kv))
        np.testing.assert_array_equal(out.get(), np.where(kv < 0, exact, np.minimum(exact, kv + 1)),
                                      err_msg=str(p))
    pq, pc = W.cdist(2, 16, 4000, seed=5)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    kq = rng.choice([-1, 0, 2, 4], size=Q.n).astype(np.int32)
    out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
    hip.cdist_batched_into(Q, C, out, -1, None, kvec=hip.DeviceBuffer.from_host(kq))
    exact = O.cross(pq, pc, threads=8)
    want = np.where(kq[:, None] < 0, exact, np.minimum(exact, kq[:, None] + 1))
    np.testing.assert_array_equal(out.get(), want)


def test_graph_replays_match_eager():
    s = hip.Stream()
    pa, pb = _mixed_pairs(3, 4000)
    A, B = hip.Packed(*pa, stream=s.ptr), hip.Packed(*pb, stream=s.ptr)
    pq, pc = W.cdist(2, 8, 3000, seed=6)
    Q, C = hip.Packed(*pq, stream=s.ptr), hip.Packed(*pc, stream=s.ptr)
    o1 = hip.DeviceArray.empty((A.n,), np.uint16)
    o2 = hip.DeviceArray.empty((Q.n, C.n), np.uint8)
    ws1, ws2 = hip.Workspace(), hip.Workspace()

    def work():
        hip.pairs_batched_into(A, B, o1, 7, True, s.ptr, ws1)
        hip.cdist_batched_into(Q, C, o2, 2, s.ptr, ws2)

    work()
    s.sync()
    eager1, eager2 = o1.get(), o2.get()
    np.testing.assert_array_equal(eager1, O.saturate(O.pairs(pa, pb, k=7, threads=8), np.uint16))
    np.testing.assert_array_equal(eager2, O.saturate(O.cross(pq, pc, k=2, threads=8), np.uint8))
    g = hip.Graph(work, s)
    for rep in range(4):
        call("bl_memset", o1.ptr, 0xAB, o1.nbytes, s.ptr)          # eager work between replays
        call("bl_memset", o2.ptr, 0xCD, o2.nbytes, s.ptr)
        s.sync()
        g.launch()
        s.sync()
        np.testing.assert_array_equal(o1.get(), eager1, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(o2.get(), eager2, err_msg=f"replay {rep}")


def test_autotuner_picks_and_caches(tmp_path, monkeypatch):
    """The tuners time their candidates (band tiles included for pairs), return
    a shape that gives the oracle's results, and answer the next call from
    their cache file."""
    monkeypatch.setenv("BITLEV_TUNE_CACHE", str(tmp_path / "tune.json"))
    pa, pb = W.pairs(3, 30000, seed=6)
    A, B = hip.Packed(*pa), hip.Packed(*pb)
    out = hip.DeviceArray.empty((A.n,), np.uint16)
    seen = []
    p = autotune.tune_pairs(A, B, out, -1, True, None, sample_pairs=8192, log=seen.append)
    assert any("'band'" in line for line in seen)
    hip.pairs_batched_into(A, B, out, -1, True, None, None, p)
    np.testing.assert_array_equal(out.get(), O.saturate(O.pairs(pa, pb, threads=8), np.uint16))
    assert autotune.tune_pairs(A, B, out, -1, True, None, log=seen.append) == p
    pq, pc = W.cdist(2, 300, 20000, seed=7)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    outc = hip.DeviceArray.empty((Q.n, C.n), np.uint8)
    pc_ = autotune.tune_cdist(Q, C, outc, 2, None, sample_queries=64)
    hip.cdist_batched_into(Q, C, outc, 2, None, None, pc_)
    np.testing.assert_array_equal(outc.get(), O.saturate(O.cross(pq, pc, k=2, threads=8), np.uint8))
    assert autotune.tune_cdist(Q, C, outc, 2, None) == pc_
    assert (tmp_path / "tune.json").exists()


def test_gather_cols_every_dtype():
    """bl_gather_cols: dst[r][i] = src[r][inv[i]] for every output dtype, a
    column count that is not a multiple of the block, and one row."""
    rng = np.random.default_rng(14)
    for rows, cols in ((5, 1000), (1, 77), (3, 70000)):
        inv = rng.permutation(cols).astype(np.int64)
        dinv = hip.DeviceBuffer.from_host(inv)
        for dt in (np.uint8, np.uint16, np.int32):
            src = rng.integers(0, 200, (rows, cols)).astype(dt)
            ds = hip.DeviceBuffer.from_host(src)
            dd = hip.DeviceArray.empty((rows, cols), dt)
            call("bl_gather_cols", ds.ptr, dinv.ptr, dd.ptr, rows, cols, hip.DTYPE_CODE[np.dtype(dt)], None)
            call("bl_device_sync")
            np.testing.assert_array_equal(dd.get(), src[:, inv])


def test_count_kernel_fast_paths_every_output():
    """The count kernel's fast loops (one threshold for all queries, byte
    lengths, aligned rows; pairs of windows for one-word queries) for every
    output dtype and the match list, on a search workload with near
    duplicates, against the oracle; an odd choice count leaves a tail window
    and unaligned rows."""
    for nc in (6000, 6003):
        pq, pc = W.cdist(2, 20, nc, seed=21)
        Q, C = hip.Packed(*pq), hip.Packed(*pc)
        exact = O.cross(pq, pc, threads=8)
        for k in (0, 2, 3):
            want = np.minimum(exact, k + 1)
            for tile, g in ((2048, 64), (1024, 8), (512, 3)):
                p = {"mode": "count", "tile": tile, "qgroup": g}
                for dt in (np.uint8, np.uint16, np.int32):
                    out = hip.DeviceArray.empty((Q.n, C.n), dt)
                    hip.cdist_batched_into(Q, C, out, k, None, None, p)
                    np.testing.assert_array_equal(out.get(), O.saturate(want, dt),
                                                  err_msg=f"nc={nc} k={k} {p} {dt}")
                qi, ci, di = hip.cdist_matches_into(Q, C, k, params=p)
                got = sorted(zip(qi.get().tolist(), ci.get().tolist(), di.get().tolist()))
                wq, wc = np.nonzero(exact <= k)
                assert got == sorted(zip(wq.tolist(), wc.tolist(), exact[wq, wc].tolist())), (nc, k, p)


def test_count_kernel_long_queries_rescans():
    """The count kernel's full-pattern rescans for queries of 129-256 and
    257-512 bytes (four- and eight-word Peq tables): near copies of each query
    (a few substitutions, insertions and deletions) reach the escalation
    stage, mixed with unrelated choices; per-query thresholds include an exact
    query (-1, run lane-strided) and thresholds above 255; dense and
    match-list output against the oracle."""
    rng = np.random.default_rng(71)
    Qs = [bytes(rng.integers(97, 101, m).astype(np.uint8)) for m in (130, 256, 257, 400, 512)]
    Cs = []
    for i in range(3000):
        q = Qs[i % len(Qs)]
        b = bytearray(q)
        if i % 4 == 3:
            b = bytearray(rng.integers(97, 101, int(rng.integers(100, 520))).astype(np.uint8))
        else:
            for _ in range(int(rng.integers(0, 12))):
                op = int(rng.integers(0, 3))
                p_ = int(rng.integers(0, len(b)))
                ch = 97 + int(rng.integers(0, 4))
                if op == 0:
                    b[p_] = ch
                elif op == 1:
                    b.insert(p_, ch)
                else:
                    del b[p_]
        Cs.append(bytes(b))
    pq, pc = O.pack(Qs), O.pack(Cs)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    for kq in ([6, 6, 6, 6, 6], [3, 9, -1, 300, 6]):
        kv = np.array(kq, np.int32)
        want = np.where(kv[:, None] < 0, exact, np.minimum(exact, kv[:, None] + 1))
        for p in ({"mode": "count", "tile": 256, "qgroup": 2}, {"mode": "count", "qgroup": 64},
                  {"mode": "count", "tile": 1024, "qgroup": 4, "waves_per_cu": 1}):
            out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
            hip.cdist_batched_into(Q, C, out, -1, None, None, p, kvec=hip.DeviceBuffer.from_host(kv))
            np.testing.assert_array_equal(out.get(), want, err_msg=f"{kq} {p}")
            if min(kq) >= 0:
                qi, ci, di = hip.cdist_matches_into(Q, C, -1, params=p, kvec=hip.DeviceBuffer.from_host(kv))
                got = sorted(zip(qi.get().tolist(), ci.get().tolist(), di.get().tolist()))
                wq, wc = np.nonzero(exact <= kv[:, None])
                assert got == sorted(zip(wq.tolist(), wc.tolist(), exact[wq, wc].tolist())), (kq, p)


def test_untuned_thresholded_pairwise_default_shape():
    """Public pairwise calls without a launch shape: thresholds 0-30 run band
    tiles from two units (the default shape), others the library default; on
    near-duplicate pairs of lengths 1-256 (where the band decides most pairs)
    plus empty and long strings, every dtype, against the oracle."""
    pa, pb = W.pairs(3, 6000, seed=12)
    A = [bytes(pa[0][pa[1][i]:pa[1][i + 1]]) for i in range(len(pa[1]) - 1)]
    B = [bytes(pb[0][pb[1][i]:pb[1][i + 1]]) for i in range(len(pb[1]) - 1)]
    A += [b"", b"x" * 300, b"ab" * 150]
    B += [b"abc", b"", b"ab" * 149 + b"a"]
    assert hip.default_pairs_params(8) is not None and hip.default_pairs_params(31) is None
    pk = O.pack(A), O.pack(B)
    for k in (0, 1, 8, 30, 31, None):
        kk = -1 if k is None else k
        want = O.pairs(*pk, k=kk, threads=8)
        for dt in (np.int32, np.uint8, np.uint16):
            got = hip.pairwise(A, B, max_dist=k, dtype=dt)
            np.testing.assert_array_equal(got, O.saturate(want, dt), err_msg=f"k={k} {dt}")


def _short_sets(seed, nq, nc, maxlen=64):
    """Queries and choices of at most `maxlen` bytes over several alphabets
    (26 letters, DNA, two letters, all bytes), with near copies of queries,
    anagrams, runs and empty strings among the choices."""
    rng = np.random.default_rng(seed)
    alphas = [np.arange(97, 123), np.frombuffer(b"ACGT", np.uint8), np.array([97, 98]),
              np.arange(256)]
    Q = []
    for i in range(nq):
        a = alphas[i % 4]
        Q.append(bytes(rng.choice(a, int(rng.integers(1, maxlen + 1))).astype(np.uint8)))
    C = []
    for i in range(nc):
        r = rng.random()
        q = bytearray(Q[int(rng.integers(nq))])
        if r < 0.3:
            for _ in range(int(rng.integers(0, 6))):
                op = int(rng.integers(0, 3))
                pos = int(rng.integers(0, len(q) + 1))
                if op == 0 and pos < len(q):
                    q[pos] = int(rng.integers(0, 256))
                elif op == 1 and len(q) < maxlen:
                    q.insert(pos, int(rng.integers(0, 256)))
                elif pos < len(q):
                    del q[pos]
            C.append(bytes(q))
        elif r < 0.4:
            x = np.frombuffer(bytes(q), np.uint8).copy()
            rng.shuffle(x)
            C.append(x.tobytes())
        elif r < 0.45:
            C.append(bytes([int(rng.integers(0, 256))]) * int(rng.integers(0, maxlen + 1)))
        else:
            a = alphas[int(rng.integers(4))]
            C.append(bytes(rng.choice(a, int(rng.integers(0, maxlen + 1))).astype(np.uint8)))
    return O.pack(Q), O.pack(C)


@pytest.mark.parametrize("nc", [5120, 5123])
def test_count_kernel_matrix_core_tiles(nc):
    """The count kernel's int8-MFMA tiles (queries and choice windows of at
    most 64 bytes: the class-count L2 bound on 32 x 32 pairs at a time) for
    thresholds up to and past where the bound can prune (k = 63, 64, 100),
    partial query groups of 5, 40 and 64, every output dtype with 16-byte
    rows and without, and the match list; the same launch with the tiles
    switched off gives the same bytes, and both equal the oracle. Every
    fifth choice window holds a 65-byte choice and falls back to the
    byte-class path."""
    pq, pc = _short_sets(31, 69, nc)
    C = O.unpack(*pc)
    for w in range(0, nc, 1280):                 # some windows not eligible for the tiles
        C[min(w + 7, nc - 1)] = b"z" * 65
    pc = O.pack(C)
    Q, Cp = hip.Packed(*pq), hip.Packed(*pc)
    exact = O.cross(pq, pc, threads=8)
    for k in (0, 1, 2, 3, 8, 30, 63, 64, 100):
        want = np.minimum(exact, k + 1)
        for g in (5, 40, 64):
            for gram in (True, False):
                p = {"mode": "count", "tile": 1024, "qgroup": g, "gram": gram}
                for dt in (np.uint8, np.uint16, np.int32):
                    out = hip.DeviceArray.empty((Q.n, Cp.n), dt)
                    hip.cdist_batched_into(Q, Cp, out, k, None, None, p)
                    np.testing.assert_array_equal(out.get(), O.saturate(want, dt),
                                                  err_msg=f"nc={nc} k={k} {p} {dt}")
            if k <= 30:
                p = {"mode": "count", "tile": 1024, "qgroup": g}
                qi, ci, di = hip.cdist_matches_into(Q, Cp, k, params=p)
                got = sorted(zip(qi.get().tolist(), ci.get().tolist(), di.get().tolist()))
                wq, wc = np.nonzero(exact <= k)
                assert got == sorted(zip(wq.tolist(), wc.tolist(), exact[wq, wc].tolist())), (k, p)


@pytest.mark.parametrize("gen,k", [(4, 8), (2, 2)])
def test_count_mode_escalation_list_sizes(gen, k):
    """Count mode lists its near matches for the band kernel in the rest of
    the workspace; with no room for a list (a workspace that ends before it),
    a list of 10 or 100 entries (overflow: the rest scanned in the count
    kernel) or the full default list, dense and match-list results equal the
    oracle. Near copies make the list long (k = 8 DNA, k = 2 letters)."""
    pq, pc = W.cdist(gen, 24, 6000, seed=11)
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    st = C.stripes()
    want = O.cross(pq, pc, k=k, threads=8)
    mq = Q.max_len
    full = call("bl_cdist_workspace_size", Q.n, mq)
    entries = min(Q.n * 2048, 1 << 20)
    core = full - 256 - 8 * entries
    assert core > 0
    for size, use_list in ((core, True), (core + 256 + 8 * 10, True), (core + 256 + 8 * 100, True),
                           (full, True), (full, False)):
        p = hip._cdist_params({"mode": "count", "tile": 1024, "qgroup": 16, "list": use_list})
        ws = hip.DeviceBuffer(size)
        out = hip.DeviceArray.empty((Q.n, C.n), np.uint8)
        call("bl_cdist_tuned", Q.d_buf.ptr, Q.d_off.ptr, Q.n, st.ref, mq, out.ptr, hip.DTYPE_CODE[out.dtype], k,
             None, ws.ptr, ws.nbytes, p, None)
        call("bl_stream_sync", None)
        np.testing.assert_array_equal(out.get(), O.saturate(want, np.uint8), err_msg=f"workspace {size}")
        cap = max(Q.n * C.n, 1)
        dq, dc, dd = (hip.DeviceArray.empty((cap,), t) for t in (np.int32, np.int64, np.int32))
        cnt = hip.DeviceBuffer(8)
        call("bl_cdist_matches", Q.d_buf.ptr, Q.d_off.ptr, Q.n, st.ref, mq, k, None, dq.ptr, dc.ptr, dd.ptr,
             cap, cnt.ptr, ws.ptr, ws.nbytes, p, None)
        n = np.zeros(1, np.uint64)
        call("bl_memcpy_dtoh", n.ctypes.data, cnt.ptr, 8, None)
        call("bl_stream_sync", None)
        n = int(n[0])
        got = sorted(zip(dq.get()[:n].tolist(), dc.get()[:n].tolist(), dd.get()[:n].tolist()))
        wq, wc = np.nonzero(want <= k)
        assert got == sorted(zip(wq.tolist(), wc.tolist(), want[wq, wc].tolist())), size
    nw = 1 if mq <= 64 else 2 if mq <= 128 else 4 if mq <= 256 else 8
    needed = 4096 + Q.n * 256 * nw * 8 + Q.n * (16 + 16 + 64)   # queues, Peq tables, profiles
    assert needed <= core < needed + 256
    with pytest.raises(ValueError):                      # below the part before the list
        ws = hip.DeviceBuffer(needed - 8)
        out = hip.DeviceArray.empty((Q.n, C.n), np.uint8)
        call("bl_cdist_tuned", Q.d_buf.ptr, Q.d_off.ptr, Q.n, st.ref, mq, out.ptr, 2, k, None, ws.ptr, ws.nbytes,
             p, None)
