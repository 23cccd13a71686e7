"""CPU-side checks of the boundary: the HIP library loads and exports every
C ABI symbol include/*.h declares, the packing glue keeps the reference's
contract (tests/test_pack.py:9-28, tests/test_phase3b.py:68-79 there), the
argument checks match bitlev/_common.py, and the host class split matches
bench/cudamyers.py:23-36 (tests/test_cuda_host.py:13-41 there). No GPU calls."""
from __future__ import annotations

import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import PKG, ROOT

import bitlev
from bitlev import _capi, _edpy, hip
from bitlev._common import check_dtype, check_k


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(bl_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_capi.LIB_PATH)          # only to read the export table
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    _capi.lib()
    assert set(names) == set(_capi.symbols()), set(names) ^ set(_capi.symbols())


def test_extension_binds_every_entry_point():
    """The package reaches the library through its C extension (bitlev._edpy,
    the reference's binding route, csrc/edpy.c:12, :231-239), one function per
    entry point; no product module loads it with ctypes."""
    mod = _capi.lib()
    assert mod is _edpy
    for n in declared_symbols():
        assert callable(getattr(_edpy, n)), n
    for f in glob.glob(os.path.join(PKG, "bitlev", "*.py")):
        if os.path.basename(f) == "workloads.py":      # synthetic-input generator, not the ABI
            continue
        assert "CDLL" not in open(f).read(), f
    with pytest.raises(TypeError):
        _edpy.bl_free("not an address")
    assert _edpy.bl_cdist_workspace_size(-1, 0) == 0
    core = 4096 + 4 * 2048 + 4 * 32 + 4 * 64          # queues, Peq tables, profiles
    assert _edpy.bl_cdist_workspace_size(4, 64) == (core + 255) // 256 * 256 + 256 + 4 * 2048 * 8


def test_class_entry_points_reject_bad_arguments():
    """Argument errors come back as BL_EINVAL before anything is launched (no
    GPU needed: every case below fails the checks)."""
    _capi.lib()
    assert _edpy.bl_cdist_class(3, 1, 1, 1, 1, 1, 1, 1, 1, 0, -1, None) == -2     # nw not in {1,2,4,8}
    assert _edpy.bl_cdist_class(1, 1, 1, 1, -1, 1, 1, 1, 1, 0, -1, None) == -2    # negative count
    assert _edpy.bl_cdist_class(1, 1, 1, 1, 1, 1, 1, 1, 1, 7, -1, None) == -2     # bad dtype
    assert _edpy.bl_cdist_class(1, 0, 1, 1, 1, 1, 1, 1, 1, 0, -1, None) == -2     # null input
    assert _edpy.bl_cdist_class(1, 1, 1, 1, 0, 1, 1, 1, 1, 0, -1, None) == 0      # nothing to do
    assert _edpy.bl_pairs_class(48, 1, 1, 1, 1, 1, 1, 1, 0, 1, -1, None) == -2    # cap


def test_library_is_gfx950_code_object():
    data = open(_capi.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    assert b"sm_" not in data.split(b"gfx950")[0][-64:]


def test_available_never_raises_without_gpu():
    assert isinstance(hip.available(), bool)
    assert isinstance(bitlev.available(), bool)


def test_device_count_is_callable():
    assert _capi.lib().bl_device_count() >= 0
    assert _capi.lib().bl_version() == 1


@pytest.mark.parametrize("S", [[], [""], ["", ""], ["ab", "", "xyz"], ["\xff\x00\x80"],
                               ["q" * 1000] * 3, [b"by", "tes"]])
def test_pack_layout(S):
    buf, off = hip.pack(S)
    enc = [s.encode("latin-1") if isinstance(s, str) else s for s in S]
    assert off.dtype == np.int64 and off.tolist() == list(np.cumsum([0] + [len(e) for e in enc]))
    assert bytes(buf[:off[-1]]) == b"".join(enc)
    assert len(buf) >= 1                         # never empty: the GPU needs a pointer


def test_pack_errors():
    with pytest.raises(ValueError):
        hip.pack(["ok", "ā"])
    with pytest.raises(TypeError):
        hip.pack(["ok", 3])
    with pytest.raises(TypeError):
        hip.pack(5)
    n, total, mx = _edpy.sizes(["ab", b"", "xyz"])
    assert (n, total, mx) == (3, 5, 3)
    b = bytearray(5)
    o = np.zeros(4, np.int64)
    assert _edpy.pack_into(["ab", b"", "xyz"], b, o) == 5 and bytes(b) == b"abxyz"
    with pytest.raises(ValueError):
        _edpy.pack_into(["abc"], bytearray(2), np.zeros(2, np.int64))


def test_pack_range_matches_pack_on_every_span():
    """The GIL-free range packer: same bytes and offsets as pack() of the
    slice, a short buffer reports the size it needs, errors as pack()'s."""
    rng = np.random.default_rng(5)
    S = []
    for i in range(300):
        s = bytes(rng.integers(0, 256, int(rng.integers(0, 90))).astype(np.uint8))
        S.append(s.decode("latin-1") if i % 2 else s)
    for seq in (S, tuple(S)):
        for lo, hi in ((0, 0), (0, 300), (17, 18), (40, 261), (299, 300)):
            buf, off = hip.pack(S[lo:hi])
            b = np.zeros(int(off[-1]) + 3, np.uint8)
            o = np.full(hi - lo + 1, -1, np.int64)
            total, mx = _edpy.pack_range(seq, lo, hi, b, o)
            assert total == off[-1] and o.tolist() == off.tolist()
            assert mx == (int(np.diff(off).max()) if hi > lo else 0)
            assert bytes(b[:total]) == bytes(buf[:total])
            if total:
                short = np.zeros(total - 1, np.uint8)
                assert _edpy.pack_range(seq, lo, hi, short, o)[0] == total
    with pytest.raises(TypeError):
        _edpy.pack_range(iter(S), 0, 1, bytearray(8), np.zeros(2, np.int64))
    with pytest.raises(IndexError):
        _edpy.pack_range(S, 299, 301, bytearray(8), np.zeros(3, np.int64))
    with pytest.raises(ValueError):
        _edpy.pack_range(S, 0, 4, bytearray(8), np.zeros(4, np.int64))
    with pytest.raises(ValueError):
        _edpy.pack_range(["ok", "ā"], 0, 2, bytearray(8), np.zeros(3, np.int64))
    with pytest.raises(TypeError):
        _edpy.pack_range(["ok", 3], 0, 2, bytearray(8), np.zeros(3, np.int64))


def test_argument_checks():
    assert check_k(None) == -1 and check_k(3) == 3 and check_k(2**40) == 2**30
    with pytest.raises(ValueError):
        check_k(-1)
    for dt in (np.int32, np.uint16, np.uint8, "uint8"):
        check_dtype(dt)
    with pytest.raises(ValueError):
        check_dtype(np.float32)
    with pytest.raises(ValueError):
        bitlev.pairwise(["a"], ["b"], method="neon")      # not a method of this package


def test_api_errors_before_any_device_work():
    with pytest.raises(ValueError):
        hip.pairwise(["a"], ["a", "b"])
    with pytest.raises(ValueError):
        hip.cdist(["a"], ["b"], max_dist=-1)
    with pytest.raises(ValueError):
        hip.cdist(["a"], ["b"], dtype=np.float32)
    with pytest.raises(ValueError):
        hip.distance("ok", "ā")
    with pytest.raises(TypeError):
        hip.distance("ok", 3)


def _off(lengths):
    o = np.zeros(len(lengths) + 1, dtype=np.int64)
    o[1:] = np.cumsum(lengths)
    return o


def test_pair_classes_use_shorter_length_and_boundaries():
    la = [0, 16, 17, 32, 33, 64, 65, 128, 129, 256, 257, 300, 5]
    lb = [9, 99, 17, 32, 400, 64, 65, 999, 129, 256, 257, 10, 300]
    groups, rest = hip.pair_classes(_off(la), _off(lb))
    got = {int(i): c for c, g in enumerate(groups) for i in g}
    for i, mm in enumerate(min(x, y) for x, y in zip(la, lb)):
        if mm > 256:
            assert i in rest and i not in got
        else:
            cap = hip.PAIR_CAPS[got[i]]
            assert mm <= cap and (got[i] == 0 or mm > hip.PAIR_CAPS[got[i] - 1])
    assert sorted(list(rest) + list(got)) == list(range(len(la)))


def test_query_classes():
    lq = [0, 1, 64, 65, 128, 129, 256, 257, 512, 513]
    groups, rest = hip.query_classes(_off(lq))
    got = {int(i): hip.CROSS_NWS[c] for c, g in enumerate(groups) for i in g}
    assert got == {0: 1, 1: 1, 2: 1, 3: 2, 4: 2, 5: 4, 6: 4, 7: 8, 8: 8}
    assert list(rest) == [9]


def test_error_mapping():
    with pytest.raises(MemoryError):
        _capi.check(_capi.BL_EALLOC, "x")
    with pytest.raises(ValueError):
        _capi.check(_capi.BL_EINVAL, "x")
    with pytest.raises(_capi.HipError):
        _capi.check(_capi.BL_EHIP, "x")
    assert _capi.check(0) == 0


def test_no_product_module_imports_the_oracle():
    for path in glob.glob(os.path.join(PKG, "bitlev", "*.py")):
        text = open(path).read()
        assert "oracle" not in text.replace("oracle.py", ""), path


def test_stripes_layout_on_host():
    """bl_stripes_layout: groups of 64 strings, each 64 x its widest string in
    dwords (at least one); the last entry is the total."""
    lens = np.array([0] * 64 + [1, 4, 5, 200] + [3] * 60 + [9], np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    g = np.empty((len(lens) + 63) // 64 + 1, np.int64)
    total = _capi.call("bl_stripes_layout", off.ctypes.data, len(lens), g.ctypes.data)
    assert g.tolist() == [0, 64, 64 + 64 * 50, 64 + 64 * 50 + 64 * 3]
    assert total == g[-1]
    bad = np.array([0, 5, 3], np.int64)
    with pytest.raises(ValueError):
        _capi.call("bl_stripes_layout", bad.ctypes.data, 2, g.ctypes.data)
    assert _capi.call("bl_stripes_layout", off.ctypes.data, 0, g.ctypes.data) == 0


@pytest.mark.parametrize("block", [64, 128, 4096])
def test_stripes_layout_sorted_on_host(block):
    """bl_stripes_layout_sorted: within each block of slots the strings are
    sorted by length, ties in string order (a stable sort, checked against
    numpy's), no string leaves its block, and the group bases are those of
    bl_stripes_layout over the strings in slot order. Lengths with a range
    above the counting sort's (>= 65536) take the comparison sort."""
    rng = np.random.default_rng(block)
    for lens in (rng.integers(0, 130, 10_000), rng.integers(0, 3, 777),
                 np.concatenate([rng.integers(0, 50, 300), [70_000], rng.integers(0, 9, 200)]),
                 np.array([5], np.int64)):
        lens = lens.astype(np.int64)
        n = len(lens)
        off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        order = np.empty(n, np.int32)
        g = np.empty((n + 63) // 64 + 1, np.int64)
        total = _capi.call("bl_stripes_layout_sorted", off.ctypes.data, n, block, order.ctypes.data,
                           g.ctypes.data)
        for b0 in range(0, n, block):
            sl = slice(b0, min(b0 + block, n))
            want = b0 + np.argsort(lens[sl], kind="stable")
            assert order[sl].tolist() == want.tolist()
        soff = np.concatenate([[0], np.cumsum(lens[order])]).astype(np.int64)
        g2 = np.empty_like(g)
        assert total == _capi.call("bl_stripes_layout", soff.ctypes.data, n, g2.ctypes.data)
        assert g.tolist() == g2.tolist()
    g = np.empty(4, np.int64)
    o = np.empty(4, np.int32)
    off = np.array([0, 1, 2, 3], np.int64)
    for bad_block in (0, 63, 100):
        with pytest.raises(ValueError):
            _capi.call("bl_stripes_layout_sorted", off.ctypes.data, 3, bad_block, o.ctypes.data,
                       g.ctypes.data)
    bad = np.array([0, 5, 3], np.int64)
    with pytest.raises(ValueError):
        _capi.call("bl_stripes_layout_sorted", bad.ctypes.data, 2, 64, o.ctypes.data, g.ctypes.data)


def test_match_list_needs_a_threshold():
    with pytest.raises(ValueError):
        hip.cdist_matches(["abc"], ["abd"], None)


@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 40000, (1 << 15) + 3])
def test_parallel_pack_matches_serial(n):
    from bitlev import ingest
    rng = np.random.default_rng(n)
    S = [bytes(rng.integers(0, 256, int(l)).astype(np.uint8)) for l in rng.integers(0, 40, n)]
    S = [s.decode("latin-1") if i % 3 == 0 else s for i, s in enumerate(S)]
    b0, o0 = _edpy.pack(S)
    for threads in (1, 3, None):
        buf, off = ingest.pack(tuple(S) if threads == 3 else S, threads=threads)
        assert off.tolist() == np.frombuffer(o0, np.int64).tolist()
        assert bytes(buf[:off[-1]]) == bytes(b0)[:off[-1]] and len(buf) >= 1
    buf, off = hip.pack(S)                       # the public packer routes long lists here
    assert off.tolist() == np.frombuffer(o0, np.int64).tolist()
    if n > 10:
        with pytest.raises(TypeError):
            ingest.pack(S[:-1] + [None])


def test_struct_bindings_follow_the_header():
    """The ctypes mirrors of the C ABI's structures list the header's fields
    in the header's order (a missing trailing field would hand the library
    whatever follows the Python object)."""
    text = open(os.path.join(ROOT, "include", "bitlev_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    for cname, py in (("bl_stripes", _capi.Stripes), ("bl_cdist_params", _capi.CdistParams),
                      ("bl_pairs_params", _capi.PairsParams)):
        body = re.search(r"typedef struct \{([^{}]*)\}\s*" + cname + ";", text).group(1)
        fields = re.findall(r"\b([a-z_0-9]+)\s*(?:\[\d+\])?\s*;", body)
        assert fields == [f[0] for f in py._fields_], (cname, fields)
