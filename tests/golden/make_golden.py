"""Write the golden vectors in tests/golden/golden.npz.

Run in a container that has the reference sources (karims/editdist-bench) at
/root/reference, after `make -C oracle ref` built them into oracle/_ref/:

    python tests/golden/make_golden.py

Inputs are seeded and generated here; every expected output comes from the
reference's own C code, compiled where it lies:
  * ed_pairs_opt / ed_pairs_k / ed_cross      csrc/myers.c:157-225
  * ref_pairs_k / ref_cross_k                 the _edpy.pairs / _edpy.cross loops
                                              (csrc/edpy.c:59-127) around the
                                              reference's static threshold functions
  * gpu_pair_distance_host / gpu_peq_distance_host / gpu_store8_host
                                              csrc/gpu_hostcheck.cpp (the reference's
                                              host build of its GPU device functions)
The case lists follow the reference's tests: tests/test_cmyers.py:13-70,
tests/test_gpu_logic.py:42-122, tests/test_cuda_gpu.py:24-76,
tests/test_lib.py:30-93, tests/test_phase3b.py:41-60.
Only data (inputs and outputs) is written; nothing of the reference is copied.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Ref, _i32p, _i64p, _ptr, _u8p, pack  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
BOUNDARY = [0, 1, 2, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 191, 192, 193,
            255, 256, 257, 512, 513]
KS = [-1, 0, 1, 2, 3, 5, 10, 17, 100]


def rand_bytes(rng, n, alphabet: bytes) -> bytes:
    a = np.frombuffer(alphabet, dtype=np.uint8)
    return a[rng.integers(0, len(a), size=n)].tobytes()


def mutate(rng, s: bytes, edits: int, alphabet: bytes) -> bytes:
    b = bytearray(s)
    for _ in range(edits):
        i = int(rng.integers(0, len(b) + 1))
        op = int(rng.integers(0, 3))
        if op == 0 and i < len(b):
            b[i] = alphabet[int(rng.integers(0, len(alphabet)))]
        elif op == 1:
            b.insert(i, alphabet[int(rng.integers(0, len(alphabet)))])
        elif i < len(b):
            del b[i]
    return bytes(b)


def pair_sets(rng):
    """Named lists of (a, b) byte pairs."""
    sets = {}
    sets["known"] = [(b"", b""), (b"", b"abc"), (b"abc", b""), (b"kitten", b"sitting"),
                     (b"flaw", b"lawn"), (b"intention", b"execution"), (b"abc", b"cba"),
                     (b"abc", b"abc"), (b"a" * 70, b"b" * 70), (b"caf\xe9", b"cafe"),
                     (b"\xff\x00", b"\x00\xff"), (b"abc", b"abd"), (b"abc", b"x")]
    edge = [(b"", b""), (b"", b"abc"), (b"abc", b""), (b"a", b"a"), (b"a", b"b"),
            (b"kitten", b"sitting"), (b"\x00", b"\x00\x00"), (b"\x00" * 5, b"\x01" * 3),
            (b"\xff" * 16, b"\xff" * 15 + b"\x00"), (bytes(range(256)), bytes(range(255, -1, -1))),
            (b"x" * 300, b"y" * 290), (b"x" * 300, b""), (b"abcab", b"abab"), (b"aaaa", b"aa"),
            (b"abab", b"baba")]
    sets["edge"] = edge + [(b, a) for a, b in edge]
    bnd = []
    for n in BOUNDARY:
        for _ in range(12):
            a = rand_bytes(rng, n, b"ACGT")
            b = rand_bytes(rng, int(rng.integers(0, n + 71)), b"ACGT")
            bnd += [(a, b), (b, a)]
    sets["boundary"] = bnd
    mixed = []
    for _ in range(1500):
        alpha = [b"AB", b"ACGT", b"abcdefghijklmnopqrstuvwxyz ", b"\x00\x01\xff",
                 bytes(range(256))][int(rng.integers(0, 5))]
        mixed.append((rand_bytes(rng, int(rng.integers(0, 301)), alpha),
                      rand_bytes(rng, int(rng.integers(0, 301)), alpha)))
    sets["mixed"] = mixed
    near = []
    for _ in range(1500):
        m = int(rng.integers(0, 300))
        a = rand_bytes(rng, m, b"ACGT\x00")
        near.append((a, mutate(rng, a, int(rng.integers(0, 15)), b"ACGT")))
    sets["near"] = near
    zero = []
    for _ in range(300):
        a = rand_bytes(rng, int(rng.integers(1, 70)), b"\x00\x01")
        zero.append((a, rand_bytes(rng, int(rng.integers(0, 90)), b"\x00\x01")))
    sets["zero"] = zero
    return sets


def cross_sets(rng):
    sets = {}
    Q = [rand_bytes(rng, int(rng.integers(0, 201)), b"ACGT") for _ in range(25)]
    C = [rand_bytes(rng, int(rng.integers(0, 201)), b"ACGT") for _ in range(31)]
    sets["acgt"] = (Q, C)
    Q = [b"", b"a" * 64, b"b" * 65, b"c" * 512, b"d" * 513, b"abc", b"q" * 600]
    C = [b"", b"aaa", b"b" * 70, b"c" * 500, b"xyz", b"q" * 590, b"d" * 520]
    sets["long_empty"] = (Q, C)
    base = [rand_bytes(rng, int(rng.integers(8, 80)), b"abcdefghij") for _ in range(12)]
    Q = base[:10] + [b""]
    C = [mutate(rng, base[int(rng.integers(0, 12))], int(rng.integers(0, 6)), b"abcdefghij")
         for _ in range(90)] + [rand_bytes(rng, int(rng.integers(0, 90)), b"abcdefghij")
                                for _ in range(40)]
    sets["text"] = (Q, C)
    Q = [rand_bytes(rng, int(rng.integers(0, 260)), bytes(range(256))) for _ in range(9)]
    C = [rand_bytes(rng, int(rng.integers(0, 260)), bytes(range(256))) for _ in range(40)]
    sets["bytes"] = (Q, C)
    return sets


def main():
    ref = Ref()
    if ref.myers is None or ref.driver is None or ref.gpucheck is None:
        sys.exit("build the reference first: make -C oracle ref")
    rng = np.random.default_rng(20261203)
    data = {}
    for name, pairs in pair_sets(rng).items():
        A, B = [p[0] for p in pairs], [p[1] for p in pairs]
        pa, pb = pack(A), pack(B)
        data[f"pairs/{name}/buf_a"], data[f"pairs/{name}/off_a"] = pa
        data[f"pairs/{name}/buf_b"], data[f"pairs/{name}/off_b"] = pb
        n = len(A)
        # csrc/myers.c entry points (trim off and on, exact)
        out = np.empty(n, np.int32)
        args = (_ptr(pa[0], _u8p), _ptr(pa[1], _i64p), _ptr(pb[0], _u8p), _ptr(pb[1], _i64p), n)
        for trim in (0, 1):
            assert ref.myers.ed_pairs_opt(*args, _ptr(out, _i32p), trim) == 0
            data[f"pairs/{name}/ed_pairs_opt/trim{trim}"] = out.copy()
        # the extension's threshold loop: every k, trim on and off
        for k in KS:
            for trim in (0, 1):
                data[f"pairs/{name}/pairs_k/k{k}/trim{trim}"] = ref.pairs(pa, pb, k, trim)
        # the reference GPU device functions on the host (-1 = class > 256)
        for k in (-1, 0, 2, 5, 17):
            for trim in (0, 1):
                v = [ref.gpucheck.gpu_pair_distance_host(a, len(a), b, len(b), trim, k)
                     for a, b in pairs]
                data[f"pairs/{name}/gpu_swar/k{k}/trim{trim}"] = np.array(v, np.int32)
            v = [ref.gpucheck.gpu_peq_distance_host(a, len(a), b, len(b), k) for a, b in pairs]
            data[f"pairs/{name}/gpu_peq/k{k}"] = np.array(v, np.int32)
    for name, (Q, C) in cross_sets(rng).items():
        pq, pc = pack(Q), pack(C)
        data[f"cross/{name}/buf_q"], data[f"cross/{name}/off_q"] = pq
        data[f"cross/{name}/buf_c"], data[f"cross/{name}/off_c"] = pc
        out = np.empty((len(Q), len(C)), np.int32)
        assert ref.myers.ed_cross(_ptr(pq[0], _u8p), _ptr(pq[1], _i64p), len(Q), _ptr(pc[0], _u8p),
                                  _ptr(pc[1], _i64p), len(C), _ptr(out, _i32p)) == 0
        data[f"cross/{name}/ed_cross"] = out.copy()
        for k in KS:
            data[f"cross/{name}/cross_k/k{k}"] = ref.cross(pq, pc, k)
    data["store8/in"] = np.array([0, 7, 254, 255, 256, 9999, 65535, 65536], np.int32)
    data["store8/out"] = np.array([ref.gpucheck.gpu_store8_host(int(d)) for d in data["store8/in"]],
                                  np.int32)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(data)} arrays, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
