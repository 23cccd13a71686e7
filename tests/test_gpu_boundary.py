"""The drop-in boundary on the GPU.

* The reference's launches swapped one for one: the ctypes stub a maintainer
  would add to the reference's bitlev/gpu.py (the one INTEGRATION.md shows)
  calls bl_pairs_class / bl_cdist_class with exactly the reference's packed
  arguments (csrc/gpu_kernels.cu:21-27, :59-65; bitlev/gpu.py:89-96,
  :121-130), and the results equal the reference's own outputs
  (tests/golden/golden.npz).
* The same entry points through the package's C extension (bitlev._edpy).
* The first benchmark workload: 65,536 pairs of 32-byte ACGT strings and
  their edited copies (p = 0.10), int32, through both forms and the public
  call, against the oracle (the reference's pairs tests,
  tests/test_cuda_gpu.py:59-76 and tests/test_lib_gpu.py:12-19 there).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import cross_of, golden_cross_sets, golden_pair_sets, pairs_of
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PAIR_CAPS = (16, 32, 64, 128, 256)
CROSS_NWS = (1, 2, 4, 8)


class Stub:
    """The maintainer's ctypes stub over libbitlev_hip.so."""

    def __init__(self):
        from bitlev import _capi
        L = ctypes.CDLL(_capi.LIB_PATH)
        vp, i64, i = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.bl_malloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.bl_free.argtypes = [vp]
        L.bl_memcpy_htod.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.bl_memcpy_dtoh.argtypes = [vp, vp, ctypes.c_size_t, vp]
        L.bl_device_sync.argtypes = []
        L.bl_pairs_class.argtypes = [i, vp, vp, vp, vp, vp, i64, vp, i, i, i, vp]
        L.bl_cdist_class.argtypes = [i, vp, vp, vp, i64, vp, vp, i64, vp, i, i, vp]
        self.L = L
        self.held = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.L.bl_malloc(ctypes.byref(p), max(int(nbytes), 1)) == 0
        self.held.append(p)
        return p

    def dev(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.bl_memcpy_htod(p, a.ctypes.data, a.nbytes, None) == 0
        return p

    def get(self, p, shape, dtype):
        h = np.empty(shape, dtype)
        assert self.L.bl_device_sync() == 0
        assert self.L.bl_memcpy_dtoh(h.ctypes.data, p, h.nbytes, None) == 0
        return h

    def free(self):
        for p in self.held:
            self.L.bl_free(p)
        self.held = []

    def pairwise(self, pa, pb, k, out8):
        """bitlev/gpu.py:71-103 with the ed_pairs_c{CAP} launch swapped."""
        n = len(pa[1]) - 1
        out = self.alloc(n * (1 if out8 else 4))
        m = np.minimum(np.diff(pa[1]), np.diff(pb[1]))
        cls = np.searchsorted(np.array(PAIR_CAPS), m)
        da, db = (self.dev(pa[0]), self.dev(pa[1])), (self.dev(pb[0]), self.dev(pb[1]))
        for ci, cap in enumerate(PAIR_CAPS):
            idx = np.nonzero(cls == ci)[0].astype(np.int32)
            if len(idx):
                rc = self.L.bl_pairs_class(cap, da[0], da[1], db[0], db[1], self.dev(idx), len(idx),
                                           out, 2 if out8 else 0, 1, k, None)
                assert rc == 0
        return self.get(out, n, np.uint8 if out8 else np.int32), cls == len(PAIR_CAPS)

    def cdist(self, pq, pc, k, out8):
        """bitlev/gpu.py:106-138 with the ed_cross_w{NW} launch swapped."""
        nq, nc = len(pq[1]) - 1, len(pc[1]) - 1
        out = self.alloc(nq * nc * (1 if out8 else 4))
        words = (np.diff(pq[1]) + 63) // 64
        cls = np.searchsorted(np.array(CROSS_NWS), words)
        dq, dc = (self.dev(pq[0]), self.dev(pq[1])), (self.dev(pc[0]), self.dev(pc[1]))
        for ci, nw in enumerate(CROSS_NWS):
            idx = np.nonzero(cls == ci)[0].astype(np.int32)
            if len(idx):
                rc = self.L.bl_cdist_class(nw, dq[0], dq[1], self.dev(idx), len(idx), dc[0], dc[1],
                                           nc, out, 2 if out8 else 0, k, None)
                assert rc == 0
        got = self.get(out, (nq, nc), np.uint8 if out8 else np.int32)
        return got, cls == len(CROSS_NWS)


@pytest.fixture(scope="module")
def stub():
    from bitlev import hip
    assert hip.available(), "no HIP device"
    s = Stub()
    yield s
    s.free()


@pytest.mark.parametrize("k", [-1, 0, 2, 10])
@pytest.mark.parametrize("out8", [False, True])
def test_stub_pairs_class_matches_golden(stub, golden, k, out8):
    for name in golden_pair_sets(golden):
        pa, pb = pairs_of(golden, name)
        got, cpu_rows = stub.pairwise(pa, pb, k, out8)
        want = golden[f"pairs/{name}/pairs_k/k{k}/trim1"]
        if out8:
            want = O.saturate(want, np.uint8)
        # rows the reference sends to its CPU core (shorter string > 256) are not
        # launched by the stub either
        np.testing.assert_array_equal(got[~cpu_rows], want[~cpu_rows], err_msg=f"{name} k={k}")
        stub.free()


@pytest.mark.parametrize("k", [-1, 0, 2, 10])
@pytest.mark.parametrize("out8", [False, True])
def test_stub_cdist_class_matches_golden(stub, golden, k, out8):
    for name in golden_cross_sets(golden):
        pq, pc = cross_of(golden, name)
        got, cpu_rows = stub.cdist(pq, pc, k, out8)
        want = golden[f"cross/{name}/cross_k/k{k}"]
        if out8:
            want = O.saturate(want, np.uint8)
        np.testing.assert_array_equal(got[~cpu_rows], want[~cpu_rows], err_msg=f"{name} k={k}")
        stub.free()


def test_class_entry_points_through_the_extension(golden):
    """bl_cdist_class and bl_cdist_class_stripes reached through bitlev._edpy
    give the reference's outputs."""
    from bitlev import _edpy, hip
    for name in golden_cross_sets(golden):
        pq, pc = cross_of(golden, name)
        Q, C = hip.Packed(*pq), hip.Packed(*pc)
        groups, rest = hip.query_classes(Q.h_off)
        for k in (-1, 2):
            want = golden[f"cross/{name}/cross_k/k{k}"]
            for use_stripes in (False, True):
                out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
                for nw, g in zip(CROSS_NWS, groups):
                    if not len(g):
                        continue
                    d_idx = hip.DeviceBuffer.from_host(g)
                    if use_stripes:
                        rc = _edpy.bl_cdist_class_stripes(nw, Q.d_buf.ptr, Q.d_off.ptr, d_idx.ptr,
                                                          len(g), C.stripes().desc, out.ptr, 0, k,
                                                          None)
                    else:
                        rc = _edpy.bl_cdist_class(nw, Q.d_buf.ptr, Q.d_off.ptr, d_idx.ptr, len(g),
                                                  C.d_buf.ptr, C.d_off.ptr, C.n, out.ptr, 0, k, None)
                    assert rc == 0
                    _edpy.bl_device_sync()
                got = out.get()
                rows = np.ones(Q.n, bool)
                rows[rest] = False
                np.testing.assert_array_equal(got[rows], want[rows], err_msg=f"{name} k={k}")


@pytest.fixture(scope="module")
def acgt32():
    from bitlev import workloads as W
    pa, pb = W.pairs(1, 65_536)
    assert len(pa[1]) - 1 == 65_536
    assert set(np.unique(pa[0]).tolist()) <= set(b"ACGT")
    assert (np.diff(pa[1]) == 32).all()
    exact = O.pairs(pa, pb, k=-1, trim=True, threads=8)
    return pa, pb, exact


@pytest.mark.parametrize("form", ["plain", "batched"])
@pytest.mark.parametrize("trim", [True, False])
@pytest.mark.parametrize("k", [-1, 3])
def test_pairs_acgt32_workload_both_forms(acgt32, form, trim, k):
    from bitlev import hip
    pa, pb, exact = acgt32
    A, B = hip.Packed(*pa), hip.Packed(*pb)
    out = hip.DeviceArray.empty((A.n,), np.int32)
    keep = hip.run_pairs(A, B, out, k, trim, None, form)
    got = out.get()
    del keep
    want = exact if k < 0 else np.minimum(exact, k + 1)
    np.testing.assert_array_equal(got, want, err_msg=f"{form} trim={trim} k={k}")
    assert exact.max() > 3 and exact.min() == 0


@pytest.mark.parametrize("max_dist", [None, 0, 3])
@pytest.mark.parametrize("dt", [np.int32, np.uint16, np.uint8])
def test_pairs_acgt32_public_call(acgt32, max_dist, dt):
    from bitlev import gpu
    pa, pb, exact = acgt32
    A = [bytes(pa[0][pa[1][i]:pa[1][i + 1]]) for i in range(8192)]
    B = [bytes(pb[0][pb[1][i]:pb[1][i + 1]]).decode("latin-1") for i in range(8192)]
    want = exact[:8192] if max_dist is None else np.minimum(exact[:8192], max_dist + 1)
    np.testing.assert_array_equal(gpu.pairwise(A, B, max_dist=max_dist, dtype=dt),
                                  O.saturate(want, np.dtype(dt)))
