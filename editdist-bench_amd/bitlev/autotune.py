"""Autotuner for the batched persistent kernels.

For a resident workload it times every candidate launch shape of the batched
pipeline on a sample of the batch (HIP events, median of a few runs), times
the fastest of each family (lane mode; direct, bucketed or band tiles) again
on the whole batch (a sample's length mix and its tail of short tiles can
rank shapes differently), walks the winner's parameters one at a time on the
whole batch, keeps the fastest and caches it per device and workload
signature in a JSON file ($BITLEV_TUNE_CACHE, default
~/.cache/bitlev/autotune.json).

cdist: lane mode (compacted lockstep passes, strided lanes, gap windows for
k >= 0, or for k <= 3 the four-choice byte filter; exact scans for exact
searches), persistent waves per CU, choices per tile. pairs: direct (one launch) or bucketed, and
for bucketed persistent waves per CU, workgroup size, text-length bucket
width, bucketing chunk; whether to trim common affixes (which never changes a
distance); then, on the two best bucketed shapes, the smallest tile (in
32-row units) that runs as a diagonal-band tile. Every candidate computes the
same results; only speed differs.
"""
from __future__ import annotations

import ctypes
import itertools
import json
import os
import threading

import numpy as np

from . import hip
from ._capi import call

CDIST_SPACE = {"mode": ["strided", "queue"], "waves_per_cu": [8, 16, 32],
               "tile": [2048, 8192, 32768]}
# the four-choice byte filter, only where it can decide (k <= 3)
CDIST_SWAR_SPACE = {"mode": ["swar"], "waves_per_cu": [8, 16, 32], "tile": [2048, 8192, 32768]}
# gap windows (any threshold): the length test on 256-choice windows, the rest staged
CDIST_GAP_SPACE = {"mode": ["gap"], "waves_per_cu": [8, 16, 32], "tile": [2048, 8192, 32768]}
# exact scans (exact searches): one workgroup per (query, choice block), the
# set in length order; only the block size matters
CDIST_EXACT_SPACE = {"mode": ["exact"], "tile": [256, 1024, 4096, 16384]}
# count windows (any threshold; choices with class profiles): a choice block
# against a group of queries, the length and class-count bounds first
CDIST_COUNT_SPACE = {"mode": ["count"], "waves_per_cu": [16, 32], "tile": [2048, 8192, 32768],
                     "qgroup": [4, 16, 64]}
PAIRS_SPACE = {"waves_per_cu": [8, 16, 32], "block": [64, 256], "bin_width": [8, 16, 32],
               "chunk": [16384, 65536, 262144], "trim": [True, False]}
# direct mode (one launch, no bucketing): only the launch shape matters
PAIRS_DIRECT_SPACE = {"mode": ["direct"], "waves_per_cu": [8, 16, 32], "block": [64, 256],
                      "trim": [True, False]}
# band tiles from this many 32-row units up (tried on the best bucketed shapes)
PAIRS_BAND_UNITS = [2, 3, 4, 5]
# lanes per pair: pairs above this many 32-row units get a whole wave each
# (the anti-diagonal wavefront) instead of one lane (8: only above 256 bytes)
PAIRS_LONG_UNITS = [4, 6]

_lock = threading.Lock()


def _cache_path():
    return os.environ.get("BITLEV_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "bitlev", "autotune.json"))


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    with _lock:
        d = _load()
        d[key] = value
        path = _cache_path()
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(d, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass


def _device_tag():
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    call("bl_device_info", name, 256, cu, None, None)
    return f"{name.value.decode().strip()}/{cu.value}cu"


def _grid(space):
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def _time(fn, stream, reps=None, budget_ms=10.0):
    """Median ms of fn() on stream, measured with HIP events: reps timed
    runs after one untimed one; by default as many as fit about budget_ms
    (3 to 15), so that shapes of sub-millisecond calls, which differ by a few
    percent, are ranked on more than three samples."""
    h0, h1 = ctypes.c_void_p(), ctypes.c_void_p()
    call("bl_event_create", h0)
    call("bl_event_create", h1)
    e0, e1 = h0.value, h1.value

    def once():
        call("bl_event_record", e0, stream)
        fn()
        call("bl_event_record", e1, stream)
        call("bl_event_sync", e1)
        v = ctypes.c_float()
        call("bl_event_elapsed_ms", e0, e1, v)
        return v.value

    try:
        fn()
        call("bl_stream_sync", stream)
        ts = [once()]
        n = reps if reps else int(min(15, max(3, budget_ms / max(ts[0], 1e-3))))
        ts += [once() for _ in range(n - 1)]
        return float(np.median(ts))
    finally:
        call("bl_event_destroy", e0)
        call("bl_event_destroy", e1)


def _family(p):
    """The structural choice a launch shape makes (what a small sample ranks
    least reliably): the lane mode for cdist; direct, bucketed or band tiles
    for pairs."""
    if p.get("mode") == "direct":
        return "direct"
    if "mode" in p:
        return p["mode"]
    return "band" if p.get("band") else "bucketed"


def _refine(res, full_fn, stream, refine, log, tag, axes):
    """Re-time on the whole batch the `refine` fastest sampled candidates of
    every family (on a sample, tiles per wave are few and the end-of-batch
    imbalance hides what band tiles or a lane mode save on the whole batch:
    on the 16.8M uneven pairs a 1M-pair sample put band tiles 2% behind,
    the whole batch 34% ahead), then walk the winner's parameters one axis
    at a time on the whole batch, keeping each change that is faster."""
    fams = {}
    for ms, p in sorted(res, key=lambda r: r[0]):
        lst = fams.setdefault(_family(p), [])
        if len(lst) < max(1, refine):
            lst.append(p)
    out = []

    def run(p):
        for ms, q in out:
            if q == p:
                return ms
        ms = _time(lambda: full_fn(p), stream)
        out.append((ms, p))
        if log:
            log(f"{tag} full {p} {ms:.3f} ms")
        return ms

    for lst in fams.values():
        for p in lst:
            run(p)
    best_ms, best = min(out, key=lambda r: r[0])
    for axis, values in axes.items():
        if axis == "band" and not best.get("band"):
            continue
        if axis in ("static", "gram", "list") and best.get("mode") != "count":
            continue
        if axis in ("long_units", "unit64") and (best.get("mode") is not None or "waves_per_cu" not in best):
            continue                          # bucketed pairs shapes only
        if axis not in best and axis not in ("band", "static", "gram", "list", "long_units", "unit64"):
            continue
        for v in values:
            q = dict(best, **{axis: v})
            ms = run(q)
            if ms < best_ms:
                best_ms, best = ms, q
    return best_ms, best


def _length_sig(packed: hip.Packed):
    L = packed.lengths
    if len(L) == 0:
        return "0"
    return f"{int(L.min())}-{int(np.median(L))}-{int(L.max())}"


def cdist_signature(Q: hip.Packed, C: hip.Packed, k: int, dtype) -> str:
    return f"cdist|{_device_tag()}|q{_length_sig(Q)}|c{_length_sig(C)}|k{k}|{np.dtype(dtype).str}"


def pairs_signature(A: hip.Packed, B: hip.Packed, k: int, dtype) -> str:
    return f"pairs|{_device_tag()}|a{_length_sig(A)}|b{_length_sig(B)}|k{k}|{np.dtype(dtype).str}"


def tune_cdist(Q: hip.Packed, C: hip.Packed, out: hip.DeviceArray, k: int = -1, stream=None,
               sample_queries: int = 256, space=None, use_cache: bool = True, log=None,
               refine: int = 3):
    """Best bl_cdist launch shape for these resident inputs (dict of params)."""
    sig = cdist_signature(Q, C, k, out.dtype)
    if use_cache:
        hit = _load().get(sig)
        if hit:
            return hit["params"]
    nq = min(Q.n, sample_queries)
    Qs = Q if nq == Q.n else hip.Packed(*_head(Q, nq))
    ws = hip.Workspace()
    res = []
    cands = list(_grid(space or CDIST_SPACE))
    if space is None and 0 <= k <= 3:
        cands += list(_grid(CDIST_SWAR_SPACE))
    if space is None and k < 0:
        cands += list(_grid(CDIST_EXACT_SPACE))
    if space is None and k >= 0:
        cands += list(_grid(CDIST_GAP_SPACE))
        if C.stripes(stream).desc.counts:
            cands += list(_grid(CDIST_COUNT_SPACE))
    for p in cands:
        ms = _time(lambda: hip.cdist_batched_into(Qs, C, out, k, stream, ws, p), stream)
        res.append((ms, p))
        if log:
            log(f"cdist {p} {ms:.3f} ms")
    if nq < Q.n and refine > 0:
        ms, best = _refine(res, lambda p: hip.cdist_batched_into(Q, C, out, k, stream, ws, p),
                           stream, 1, log, "cdist",
                           {"waves_per_cu": [8, 16, 32], "tile": [1024, 2048, 4096, 8192, 32768],
                            "qgroup": [16, 32, 64], "static": [False, True],
                            "gram": [False, True], "list": [False, True]})
    else:
        ms, best = min(res, key=lambda r: r[0])
    _store(sig, {"params": best, "ms": ms, "sample_queries": nq})
    return best


def tune_pairs(A: hip.Packed, B: hip.Packed, out: hip.DeviceArray, k: int = -1, trim=True,
               stream=None, sample_pairs: int = 1 << 20, space=None, use_cache: bool = True,
               log=None, refine: int = 3):
    """Best bl_pairs launch shape for these resident inputs (dict of params)."""
    sig = pairs_signature(A, B, k, out.dtype)
    if use_cache:
        hit = _load().get(sig)
        if hit:
            return hit["params"]
    n = min(A.n, sample_pairs)
    As = A if n == A.n else hip.Packed(*_head(A, n))
    Bs = B if n == B.n else hip.Packed(*_head(B, n))
    ws = hip.Workspace()
    res = []
    cands = list(_grid(space)) if space else list(_grid(PAIRS_DIRECT_SPACE)) + list(_grid(PAIRS_SPACE))
    if not trim:                       # trimming is off anyway: drop the duplicate shapes
        cands = [p for p in cands if p.get("trim", True)]
    def run(p):
        ms = _time(lambda: hip.pairs_batched_into(As, Bs, out, k, trim, stream, ws, p), stream)
        res.append((ms, p))
        if log:
            log(f"pairs {p} {ms:.3f} ms")

    for p in cands:
        run(p)
    if space is None:
        bucketed = sorted((r for r in res if r[1].get("mode") != "direct"), key=lambda r: r[0])
        for _, p in bucketed[:2]:
            for b in PAIRS_BAND_UNITS:
                run(dict(p, band=b))
        # lanes per pair (a whole wave for pairs above this many units) and
        # words per lane (64-row words only) on the best bucketed shape
        best_b = min((r for r in res if r[1].get("mode") != "direct"), key=lambda r: r[0])[1]
        for lu in PAIRS_LONG_UNITS:
            run(dict(best_b, long_units=lu))
        run(dict(best_b, unit64=True))
    if n < A.n and refine > 0:
        ms, best = _refine(res, lambda p: hip.pairs_batched_into(A, B, out, k, trim, stream, ws, p),
                           stream, 2, log, "pairs",
                           {"band": PAIRS_BAND_UNITS, "waves_per_cu": [8, 16, 32],
                            "chunk": [16384, 65536, 262144], "bin_width": [8, 16, 32],
                            "block": [64, 256], "long_units": PAIRS_LONG_UNITS + [8],
                            "unit64": [False, True]})
    else:
        ms, best = min(res, key=lambda r: r[0])
    _store(sig, {"params": best, "ms": ms, "sample_pairs": n})
    return best


def _head(P: hip.Packed, n: int):
    """Host copy of the first n strings of a resident set."""
    off = P.h_off[: n + 1].copy()
    buf = np.empty(max(int(off[-1]), 1), np.uint8)
    if off[-1]:
        call("bl_memcpy_dtoh", buf.ctypes.data, P.d_buf.ptr, int(off[-1]), None)
        call("bl_stream_sync", None)
    return buf, off
