"""CPU oracle for bitlev's Myers/Hyyro edit-distance path.

TEST INFRASTRUCTURE, NOT PRODUCT. Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module. The shipped backend never
imports it and has no CPU fallback.

Three independent checkers live here:
  * liboracle.so  -- the repo's C restatement (oracle/myers_oracle.c) of the
    reference core csrc/myers.c:22-225 and the extension loops csrc/edpy.c:59-127;
  * oracle/_ref/  -- the reference's own C sources compiled where they lie
    (oracle/Makefile `ref`): libref_myers.so (csrc/myers.c), libref_driver.so
    (our loops around its static threshold functions) and libref_gpucheck.so
    (csrc/gpu_hostcheck.cpp, the reference's host build of its GPU functions);
  * dp() / bigint_myers() below -- pure Python, for tiny inputs only
    (bench/reference.py:16-28 and :31-67).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_llp = ctypes.POINTER(ctypes.c_longlong)

DTYPE_CODES = {np.dtype(np.int32): 0, np.dtype(np.uint16): 1, np.dtype(np.uint8): 2}


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


# ----------------------------------------------------------------- packing --
def pack(strings) -> tuple[np.ndarray, np.ndarray]:
    """str (chars <= U+00FF) / bytes -> (uint8 buffer, int64 offsets[n+1]).

    Same layout as the reference's _edpy.pack (csrc/edpy.c:200-224): one
    concatenated buffer, never empty (a 1-byte dummy when all strings are)."""
    enc = [s.encode("latin-1") if isinstance(s, str) else bytes(s) for s in strings]
    off = np.zeros(len(enc) + 1, dtype=np.int64)
    if enc:
        np.cumsum([len(e) for e in enc], out=off[1:])
    buf = np.frombuffer(b"".join(enc) or b"\0", dtype=np.uint8).copy()
    return buf, off


def unpack(buf: np.ndarray, off: np.ndarray) -> list[bytes]:
    return [bytes(buf[off[i]:off[i + 1]]) for i in range(len(off) - 1)]


# ----------------------------------------------------------------- liboracle --
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_pairs.argtypes = [_u8p, _i64p, _u8p, _i64p, ctypes.c_int64, _i32p,
                                ctypes.c_int, ctypes.c_int, _llp, ctypes.c_int]
        L.orc_cross.argtypes = [_u8p, _i64p, ctypes.c_int64, _u8p, _i64p, ctypes.c_int64,
                                _i32p, ctypes.c_int, _llp, ctypes.c_int]
        L.orc_dp.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.orc_dp_pairs.argtypes = [_u8p, _i64p, _u8p, _i64p, ctypes.c_int64, _i32p, ctypes.c_int]
        L.orc_saturate.argtypes = [_i32p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
        L.orc_set_count_rule.argtypes = [ctypes.c_int]
        L.orc_set_count_rule.restype = None
        for f in (L.orc_pairs, L.orc_cross, L.orc_dp, L.orc_dp_pairs, L.orc_saturate):
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc, what):
    if rc == -1:
        raise MemoryError(f"{what}: allocation failed")
    if rc != 0:
        raise ValueError(f"{what}: invalid argument ({rc})")


def pairs(pa, pb, k: int = -1, trim: bool = True, threads: int = 1, cols: bool = False):
    """int32 distances of packed pairs (min(d, k+1) for k >= 0); optionally the
    per-pair count of text columns the reference's exit rule evaluates."""
    (ba, oa), (bb, ob) = pa, pb
    n = len(oa) - 1
    if len(ob) - 1 != n:
        raise ValueError("length mismatch")
    out = np.empty(n, dtype=np.int32)
    c = np.zeros(n, dtype=np.int64) if cols else None
    ba, bb = np.ascontiguousarray(ba), np.ascontiguousarray(bb)
    oa, ob = np.ascontiguousarray(oa, np.int64), np.ascontiguousarray(ob, np.int64)
    _check(lib().orc_pairs(_ptr(ba, _u8p), _ptr(oa, _i64p), _ptr(bb, _u8p), _ptr(ob, _i64p), n,
                           _ptr(out, _i32p), int(trim), int(k), _ptr(c, _llp), int(threads)),
           "orc_pairs")
    return (out, c) if cols else out


def count_rule(rule: str):
    """What cols=True counts per pair: "reference" (columns of the bottom-row
    exit test, csrc/myers.c:117, :127), "diagonal" (columns the final-diagonal
    bound needs) or "staged" (DP cells of the row-staged scan under the
    final-diagonal bound: 32 rows, then 64, then all rows)."""
    lib().orc_set_count_rule({"reference": 0, "diagonal": 1, "staged": 2}[rule])


def cross(pq, pc, k: int = -1, threads: int = 1, cols: bool = False):
    """Row-major nq x nc int32 matrix (no trim, the query is the pattern)."""
    (bq, oq), (bc, oc) = pq, pc
    nq, nc = len(oq) - 1, len(oc) - 1
    out = np.empty((nq, nc), dtype=np.int32)
    c = np.zeros((nq, nc), dtype=np.int64) if cols else None
    bq, bc = np.ascontiguousarray(bq), np.ascontiguousarray(bc)
    oq, oc = np.ascontiguousarray(oq, np.int64), np.ascontiguousarray(oc, np.int64)
    _check(lib().orc_cross(_ptr(bq, _u8p), _ptr(oq, _i64p), nq, _ptr(bc, _u8p), _ptr(oc, _i64p),
                           nc, _ptr(out, _i32p), int(k), _ptr(c, _llp), int(threads)),
           "orc_cross")
    return (out, c) if cols else out


def dp_pairs(pa, pb, k: int = -1) -> np.ndarray:
    """Wagner-Fischer over packed pairs (C, independent of the bit-parallel code)."""
    (ba, oa), (bb, ob) = pa, pb
    n = len(oa) - 1
    out = np.empty(n, dtype=np.int32)
    _check(lib().orc_dp_pairs(_ptr(np.ascontiguousarray(ba), _u8p), _ptr(oa, _i64p),
                              _ptr(np.ascontiguousarray(bb), _u8p), _ptr(ob, _i64p), n,
                              _ptr(out, _i32p), int(k)), "orc_dp_pairs")
    return out


def saturate(d: np.ndarray, dtype) -> np.ndarray:
    """int32 -> int32/uint16/uint8, narrow types saturating at their maximum."""
    dt = np.dtype(dtype)
    d = np.ascontiguousarray(d, dtype=np.int32)
    out = np.empty(d.shape, dtype=dt)
    _check(lib().orc_saturate(_ptr(d, _i32p), d.size, DTYPE_CODES[dt],
                              out.ctypes.data_as(ctypes.c_void_p)), "orc_saturate")
    return out


# ----------------------------------------------------------------- pure Python --
def dp(a: bytes, b: bytes) -> int:
    """Wagner-Fischer in pure Python (bench/reference.py:16-28). Tiny inputs only."""
    if len(a) < len(b):
        a, b = b, a
    prev = list(range(len(b) + 1))
    for i, ca in enumerate(a, 1):
        cur = [i] + [0] * len(b)
        for j, cb in enumerate(b, 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (ca != cb))
        prev = cur
    return prev[-1]


def bigint_myers(a: bytes, b: bytes) -> int:
    """Bit-parallel recurrence on unbounded Python ints (bench/reference.py:31-67):
    no word boundaries at all, so it checks the multi-word carry independently."""
    m = len(a)
    if m == 0:
        return len(b)
    mask = (1 << m) - 1
    top = 1 << (m - 1)
    peq: dict[int, int] = {}
    for i, c in enumerate(a):
        peq[c] = peq.get(c, 0) | (1 << i)
    vp, vn, score = mask, 0, m
    for c in b:
        eq = peq.get(c, 0)
        xv = eq | vn
        xh = ((((eq & vp) + vp) & mask) ^ vp) | eq
        hp = vn | (~(xh | vp) & mask)
        hn = vp & xh
        score += 1 if hp & top else (-1 if hn & top else 0)
        hp = ((hp << 1) | 1) & mask
        hn = (hn << 1) & mask
        vp = hn | (~(xv | hp) & mask)
        vn = hp & xv
    return score


# ----------------------------------------------------------------- reference builds --
class Ref:
    """The reference C core built by oracle/Makefile `ref` (None if absent)."""

    def __init__(self):
        self.myers = self.driver = self.gpucheck = None
        p = os.path.join(REF_DIR, "libref_myers.so")
        if os.path.exists(p):
            L = ctypes.CDLL(p)
            L.ed_pairs_opt.argtypes = [_u8p, _i64p, _u8p, _i64p, ctypes.c_int64, _i32p, ctypes.c_int]
            L.ed_pairs_k.argtypes = [_u8p, _i64p, _u8p, _i64p, ctypes.c_int64, _i32p,
                                     ctypes.c_int, ctypes.c_int]
            L.ed_cross.argtypes = [_u8p, _i64p, ctypes.c_int64, _u8p, _i64p, ctypes.c_int64, _i32p]
            self.myers = L
        p = os.path.join(REF_DIR, "libref_driver.so")
        if os.path.exists(p):
            L = ctypes.CDLL(p)
            L.ref_pairs_k.argtypes = [_u8p, _i64p, _u8p, _i64p, ctypes.c_int64, _i32p,
                                      ctypes.c_int, ctypes.c_int]
            L.ref_pairs_k_mt.argtypes = L.ref_pairs_k.argtypes + [ctypes.c_int]
            L.ref_cross_k.argtypes = [_u8p, _i64p, ctypes.c_int64, _u8p, _i64p, ctypes.c_int64,
                                      _i32p, ctypes.c_int]
            L.ref_cross_k_mt.argtypes = L.ref_cross_k.argtypes + [ctypes.c_int]
            self.driver = L
        p = os.path.join(REF_DIR, "libref_gpucheck.so")
        if os.path.exists(p):
            L = ctypes.CDLL(p)
            L.gpu_pair_distance_host.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]
            L.gpu_peq_distance_host.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                                ctypes.c_int, ctypes.c_int]
            L.gpu_store8_host.argtypes = [ctypes.c_int]
            self.gpucheck = L

    def pairs(self, pa, pb, k=-1, trim=True, threads=1):
        (ba, oa), (bb, ob) = pa, pb
        n = len(oa) - 1
        out = np.empty(n, dtype=np.int32)
        args = (_ptr(ba, _u8p), _ptr(oa, _i64p), _ptr(bb, _u8p), _ptr(ob, _i64p), n,
                _ptr(out, _i32p), int(trim), int(k))
        rc = (self.driver.ref_pairs_k(*args) if threads == 1
              else self.driver.ref_pairs_k_mt(*args, int(threads)))
        _check(rc, "ref_pairs_k")
        return out

    def cross(self, pq, pc, k=-1, threads=1):
        (bq, oq), (bc, oc) = pq, pc
        nq, nc = len(oq) - 1, len(oc) - 1
        out = np.empty((nq, nc), dtype=np.int32)
        args = (_ptr(bq, _u8p), _ptr(oq, _i64p), nq, _ptr(bc, _u8p), _ptr(oc, _i64p), nc,
                _ptr(out, _i32p), int(k))
        rc = (self.driver.ref_cross_k(*args) if threads == 1
              else self.driver.ref_cross_k_mt(*args, int(threads)))
        _check(rc, "ref_cross_k")
        return out


def effective_cpus() -> dict:
    """Logical, affinity and cgroup-quota CPU counts; `effective` is the
    smaller of the last two (method of bench/sysload.py:19-41)."""
    logical = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = logical
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, p = f.read().split()
            if q != "max":
                quota = max(1, int(int(q) / int(p)))
    except (OSError, ValueError):
        pass
    eff = min(affinity, quota) if quota else affinity
    return {"logical": logical, "affinity": affinity, "quota": quota, "effective": eff}
