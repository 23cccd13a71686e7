"""Seeded synthetic inputs for the benchmark workloads (libbl_workloads.so).

Each workload returns packed host sets (uint8 buffer, int64 offsets). The
generator is counter-based per string, so results do not depend on the thread
count. Definitions are in csrc/workloads.c; sizes and parameters of the five
benchmark configurations are in WORKLOADS below.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

SEED = 20261203

# name: (generator id, kind, sizes, max_dist, dtype)
WORKLOADS = {
    "pairs-L32-acgt": dict(gen=1, kind="pairs", n=65_536, k=None, dtype="int32"),
    "cdist-L64-k2": dict(gen=2, kind="cdist", nq=1_024, nc=1_000_000, k=2, dtype="uint8"),
    "pairs-uneven-1-256": dict(gen=3, kind="pairs", n=16_777_216, k=None, dtype="uint16"),
    "cdist-L200-256-k8": dict(gen=4, kind="cdist", nq=256, nc=4_194_304, k=8, dtype="uint8"),
    # exact searches (no max_dist, int32): the shape of the reference's published
    # GPU cdist figures, and one serving batch against a resident set
    "cdist-L64-exact": dict(gen=2, kind="cdist", nq=1_024, nc=1_000_000, k=None, dtype="int32"),
    "cdist-serving-exact": dict(gen=5, kind="cdist", nq=64, nc=1_000_000, k=None, dtype="int32"),
}


# Serving mix: resident choice sets (generator 5: lengths 16-128, 'a'-'z');
# requests half pairwise (generator 6: a of length 8-256, b = a edited at a
# per-pair rate 5-30%) and half cdist (1-64 queries of generator 5 against one
# set); max_dist uniform over {0, 1, 2, 3, 4, none}, dtype over
# {uint8, uint16, int32}.
SERVING = dict(sets=4, set_size=1_000_000, max_pairs=100_000, max_queries=64,
               ks=(0, 1, 2, 3, 4, None), dtypes=("uint8", "uint16", "int32"))


def serving_sets(seed: int = SEED, n_sets=None, set_size=None, threads=None):
    """The resident choice sets of the serving workload -> [(buf, off)]."""
    n_sets = n_sets or SERVING["sets"]
    set_size = set_size or SERVING["set_size"]
    return [generate(5, set_size, seed, 100 + s, 0, threads) for s in range(n_sets)]


def serving_requests(n: int, seed: int = SEED, n_sets=None, threads=None):
    """n request descriptions of the serving mix, each a dict with kind,
    k (None = exact), dtype and the packed payload; cdist requests name a set
    index."""
    n_sets = n_sets or SERVING["sets"]
    rng = np.random.default_rng(seed ^ 0x5E21)
    out = []
    for i in range(n):
        k = SERVING["ks"][int(rng.integers(len(SERVING["ks"])))]
        dt = SERVING["dtypes"][int(rng.integers(len(SERVING["dtypes"])))]
        rs = seed + 7919 * (i + 1)
        if rng.random() < 0.5:
            npairs = int(rng.integers(1, SERVING["max_pairs"] + 1))
            out.append(dict(kind="pairwise", k=k, dtype=dt,
                            payload=(generate(6, npairs, rs, 0, 0, threads),
                                     generate(6, npairs, rs, 1, 0, threads))))
        else:
            nq = int(rng.integers(1, SERVING["max_queries"] + 1))
            out.append(dict(kind="cdist", k=k, dtype=dt, set=int(rng.integers(n_sets)),
                            payload=generate(5, nq, rs, 200, 0, threads)))
    return out


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(os.path.join(HERE, "libbl_workloads.so"))
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.blw_lengths.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int64,
                                  ctypes.c_int64, i64p, ctypes.c_int]
        L.blw_fill.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, ctypes.c_int64,
                               ctypes.c_int64, i64p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int]
        _lib = L
    return _lib


def _threads(threads):
    if threads:
        return int(threads)
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


def generate(gen: int, n: int, seed: int = SEED, set_: int = 0, nq: int = 0, threads=None):
    """Strings [0, n) of one set of generator `gen` -> (buf, off)."""
    t = _threads(threads)
    lens = np.empty(n, np.int64)
    i64p = ctypes.POINTER(ctypes.c_int64)
    rc = lib().blw_lengths(gen, seed, set_, n, nq, lens.ctypes.data_as(i64p), t)
    if rc:
        raise ValueError(f"workload generator {gen}: error {rc}")
    off = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    buf = np.empty(max(int(off[-1]), 1), np.uint8)
    rc = lib().blw_fill(gen, seed, set_, n, nq, off.ctypes.data_as(i64p),
                        buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), t)
    if rc:
        raise ValueError(f"workload generator {gen}: error {rc}")
    return buf, off


def pairs(gen: int, n: int, seed: int = SEED, threads=None):
    """(a, b) packed sets of a pairs workload."""
    return generate(gen, n, seed, 0, 0, threads), generate(gen, n, seed, 1, 0, threads)


def cdist(gen: int, nq: int, nc: int, seed: int = SEED, threads=None):
    """(queries, choices) packed sets of a cdist workload."""
    return generate(gen, nq, seed, 0, nq, threads), generate(gen, nc, seed, 1, nq, threads)


def subset(packed, idx):
    """Host (buf, off) of a subset of a packed set, in the order of idx."""
    buf, off = packed
    idx = np.asarray(idx, np.int64)
    lens = off[idx + 1] - off[idx]
    o = np.zeros(len(idx) + 1, np.int64)
    np.cumsum(lens, out=o[1:])
    out = np.empty(max(int(o[-1]), 1), np.uint8)
    for j, i in enumerate(idx):
        out[o[j]:o[j + 1]] = buf[off[i]:off[i + 1]]
    return out, o


def head(packed, n):
    """The first n strings of a packed set (a view of the buffer)."""
    buf, off = packed
    return buf[: max(int(off[n]), 1)], off[: n + 1].copy()
