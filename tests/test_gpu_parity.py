"""GPU parity: the HIP kernels (through the C ABI) against the golden vectors
and the oracle, bit for bit. Mirrors the reference's GPU tests
(tests/test_cuda_gpu.py:24-76, tests/test_lib_gpu.py:12-35) with the CPU core
replaced by the oracle and the reference's own outputs (tests/golden)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import cross_of, golden_cross_sets, golden_pair_sets, pairs_of
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KS = [-1, 0, 1, 2, 3, 5, 10, 17, 100]
DTS = [np.int32, np.uint16, np.uint8]


@pytest.fixture(scope="module")
def hip():
    from bitlev import hip as h
    assert h.available(), "no HIP device"
    return h


FORMS = ("plain", "batched")


def gpu_pairs(hip, pa, pb, k=-1, trim=True, dtype=np.int32, form=None):
    A, B = hip.Packed(*pa), hip.Packed(*pb)
    out = hip.DeviceArray.empty((A.n,), dtype)
    keep = hip.run_pairs(A, B, out, k, trim, None, form)
    res = out.get()
    del keep
    return res


def gpu_cross(hip, pq, pc, k=-1, dtype=np.int32, form=None):
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    out = hip.DeviceArray.empty((Q.n, C.n), dtype)
    keep = hip.run_cdist(Q, C, out, k, None, form)
    res = out.get()
    del keep
    return res


@pytest.mark.parametrize("form", FORMS)
def test_golden_pairs_every_k_trim_dtype(hip, golden, form):
    for name in golden_pair_sets(golden):
        pa, pb = pairs_of(golden, name)
        print(f"pairs golden set {name} ({form})", flush=True)
        for k in KS:
            for trim in (0, 1):
                want = golden[f"pairs/{name}/pairs_k/k{k}/trim{trim}"]
                for dt in DTS:
                    got = gpu_pairs(hip, pa, pb, k, bool(trim), dt, form)
                    np.testing.assert_array_equal(got, O.saturate(want, dt),
                                                  err_msg=f"{name} k={k} trim={trim} {dt}")


@pytest.mark.parametrize("form", FORMS)
def test_golden_cross_every_k_dtype(hip, golden, form):
    for name in golden_cross_sets(golden):
        pq, pc = cross_of(golden, name)
        print(f"cross golden set {name} ({form})", flush=True)
        np.testing.assert_array_equal(gpu_cross(hip, pq, pc, form=form),
                                      golden[f"cross/{name}/ed_cross"])
        for k in KS:
            want = golden[f"cross/{name}/cross_k/k{k}"]
            for dt in DTS:
                np.testing.assert_array_equal(gpu_cross(hip, pq, pc, k, dt, form), O.saturate(want, dt),
                                              err_msg=f"{name} k={k} {dt}")


def test_known_answers_and_shapes(hip):
    assert hip.distance("kitten", "sitting") == 3
    assert hip.distance("", "abc") == 3
    assert hip.distance(b"abc", b"abd") == 1
    assert hip.distance("kitten", "sitting", max_dist=1) == 2
    assert hip.pairwise(["caf\xe9", "\xff\x00"], ["cafe", "\x00\xff"]).tolist() == [1, 2]
    r = hip.cdist(["abc"], ["abd", "x"], return_device=True)
    assert isinstance(r, hip.DeviceArray) and r.get().tolist() == [[1, 3]]
    assert hip.pairwise([], []).shape == (0,)
    assert hip.cdist([], ["a"]).shape == (0, 1)
    assert hip.cdist(["a"], []).shape == (1, 0)
    assert hip.pairwise([], [], return_device=True).shape == (0,)
    r8 = hip.pairwise(["x" * 300], [""], dtype=np.uint8)
    assert r8.dtype == np.uint8 and r8[0] == 255


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("m", [1, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257])
def test_class_boundaries_mixed_batch(hip, m, form):
    rng = np.random.default_rng(m)
    A = [bytes(rng.integers(0, 256, int(rng.integers(1, m + 1))).astype(np.uint8)) for _ in range(300)]
    B = [bytes(rng.integers(0, 256, int(rng.integers(0, m + 30))).astype(np.uint8)) for _ in range(300)]
    pa, pb = O.pack(A), O.pack(B)
    np.testing.assert_array_equal(gpu_pairs(hip, pa, pb, form=form), O.pairs(pa, pb))
    np.testing.assert_array_equal(gpu_pairs(hip, pa, pb, trim=False, form=form), O.pairs(pa, pb))


@pytest.mark.parametrize("form", FORMS)
def test_long_strings_on_gpu(hip, form):
    """The reference's CPU-fallback cases run on the wavefront kernel here,
    including multi-stripe patterns (> 4096 bytes)."""
    rng = np.random.default_rng(12)
    A = [b"x" * 300, b"a" * 257, bytes(rng.integers(0, 4, 700).astype(np.uint8)),
         bytes(rng.integers(0, 4, 5000).astype(np.uint8)), b"q" * 9000]
    B = [b"y" * 290, b"a" * 300, bytes(rng.integers(0, 4, 650).astype(np.uint8)),
         bytes(rng.integers(0, 4, 5100).astype(np.uint8)), b"q" * 8990 + b"r" * 3]
    pa, pb = O.pack(A), O.pack(B)
    for k in (-1, 0, 10, 100):
        np.testing.assert_array_equal(gpu_pairs(hip, pa, pb, k, form=form), O.pairs(pa, pb, k=k))
    Q = [b"", b"a" * 64, b"b" * 65, b"c" * 512, b"d" * 513, b"abc", b"q" * 600, b"z" * 4500]
    C = [b"", b"aaa", b"b" * 70, b"c" * 500, b"xyz", b"q" * 590, b"z" * 4400 + b"y" * 90]
    pq, pc = O.pack(Q), O.pack(C)
    for k in (-1, 0, 2, 10):
        for dt in DTS:
            np.testing.assert_array_equal(gpu_cross(hip, pq, pc, k, dt, form),
                                          O.saturate(O.cross(pq, pc, k=k), dt))


@pytest.mark.parametrize("form", FORMS)
def test_random_batch_against_oracle(hip, form):
    rng = np.random.default_rng(99)
    A, B = [], []
    for _ in range(20000):
        alpha = int(rng.choice([2, 4, 26, 256]))
        la = int(rng.integers(0, 300))
        a = rng.integers(0, alpha, la).astype(np.uint8)
        if rng.random() < 0.5:
            b = a.copy()
            for _ in range(int(rng.integers(0, 12))):
                if len(b):
                    b[int(rng.integers(0, len(b)))] = int(rng.integers(0, alpha))
            b = b[: max(0, len(b) - int(rng.integers(0, 5)))]
        else:
            b = rng.integers(0, alpha, int(rng.integers(0, 300))).astype(np.uint8)
        A.append(bytes(a))
        B.append(bytes(b))
    pa, pb = O.pack(A), O.pack(B)
    for k in (-1, 0, 3, 25):
        np.testing.assert_array_equal(gpu_pairs(hip, pa, pb, k, form=form),
                                      O.pairs(pa, pb, k=k, threads=8))


@pytest.mark.parametrize("form", FORMS)
def test_cdist_search_shape_against_oracle(hip, form):
    """A small cut of the L64 search workload: 'a'-'z', 30% near-duplicates,
    max_dist = 2, uint8, more than 64 blocks of choices."""
    rng = np.random.default_rng(5)
    Q = rng.integers(97, 123, (40, 64)).astype(np.uint8)
    C = rng.integers(97, 123, (20000, 64)).astype(np.uint8)
    near = rng.random(len(C)) < 0.3
    src = rng.integers(0, len(Q), len(C))
    C[near] = Q[src[near]]
    for i in np.nonzero(near)[0]:
        pos = rng.choice(64, int(rng.integers(0, 9)), replace=False)
        C[i, pos] = rng.integers(97, 123, len(pos))
    pq = (Q.ravel(), np.arange(0, Q.size + 1, 64, dtype=np.int64))
    pc = (C.ravel(), np.arange(0, C.size + 1, 64, dtype=np.int64))
    for k, dt in ((2, np.uint8), (-1, np.int32), (5, np.uint16)):
        want = O.saturate(O.cross(pq, pc, k=k, threads=8), dt)
        np.testing.assert_array_equal(gpu_cross(hip, pq, pc, k, dt, form), want)


@pytest.mark.parametrize("k", [None, 0, 3])
@pytest.mark.parametrize("dt", [np.int32, np.uint8, np.uint16])
def test_public_api_matches_oracle(hip, k, dt):
    rng = np.random.default_rng(3)
    A = ["".join(chr(c) for c in rng.integers(0, 256, int(rng.integers(0, 100)))) for _ in range(500)]
    B = ["".join(chr(c) for c in rng.integers(0, 256, int(rng.integers(0, 100)))) for _ in range(500)]
    A += ["x" * 300, ""]
    B += ["y" * 290, "abc"]
    kk = -1 if k is None else k
    want = O.saturate(O.pairs(O.pack(A), O.pack(B), k=kk), dt)
    got = hip.pairwise(A, B, max_dist=k, dtype=dt)
    assert got.dtype == np.dtype(dt)
    np.testing.assert_array_equal(got, want)
    Q = A[:50] + ["", "q" * 600]
    got = hip.cdist(Q, B[:700], max_dist=k, dtype=dt)
    np.testing.assert_array_equal(got, O.saturate(O.cross(O.pack(Q), O.pack(B[:700]), k=kk), dt))


def test_dlpack_export(hip):
    r = hip.pairwise(["abc", "kitten"], ["abd", "sitting"], return_device=True)
    assert r.__dlpack_device__()[0] == 10
    cap = r.__dlpack__()
    assert type(cap).__name__ == "PyCapsule"
