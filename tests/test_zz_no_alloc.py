"""bl_cdist_noalloc (the search with its workspace in a static device array)
against bl_cdist_tuned with a caller workspace, through the C ABI: the same
bytes on two calls in a row, the second with fewer and shorter queries and
fewer choices, in the library's default shape and the count kernel's; a
request above the static scratch is refused before anything runs."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sets(seed, nq, nc, qlen, clen):
    rng = np.random.default_rng(seed)
    Q = [bytes(rng.integers(97, 101, int(rng.integers(qlen[0], qlen[1] + 1))).astype(np.uint8))
         for _ in range(nq)]
    C = [bytes(rng.integers(97, 101, int(rng.integers(clen[0], clen[1] + 1))).astype(np.uint8))
         for _ in range(nc)]
    for i in range(0, nc, 7):                                 # near copies of queries
        q = bytearray(Q[i % nq])
        if q:
            q[len(q) // 2] = 120
        C[i] = bytes(q)
    return O.pack(Q), O.pack(C)


def _run(hip, entry, pq, pc, k, dtype, params):
    from bitlev import _capi, _edpy
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    st = C.stripes()
    out = hip.DeviceArray.empty((Q.n, C.n), dtype)
    _capi.call("bl_memset", out.ptr, 0xA5, out.nbytes, None)
    mq = Q.max_len
    dt = {np.dtype(np.int32): 0, np.dtype(np.uint16): 1, np.dtype(np.uint8): 2}[np.dtype(dtype)]
    prm = hip._cdist_params(params)
    if entry == "noalloc":
        rc = _edpy.bl_cdist_noalloc(Q.d_buf.ptr, Q.d_off.ptr, Q.n, st.desc, mq, out.ptr, dt, k, None,
                                    prm, None)
    else:
        ws = hip.DeviceBuffer(_edpy.bl_cdist_workspace_size(Q.n, mq))
        rc = _edpy.bl_cdist_tuned(Q.d_buf.ptr, Q.d_off.ptr, Q.n, st.desc, mq, out.ptr, dt, k, None,
                                  ws.ptr, ws.nbytes, prm, None)
    _capi.check(rc, entry)
    _edpy.bl_device_sync()
    return out.get()


@pytest.mark.parametrize("params", [None, {"mode": "count", "waves_per_cu": 16, "tile": 1024,
                                           "qgroup": 16}])
@pytest.mark.parametrize("k,dtype", [(-1, np.int32), (3, np.uint8), (40, np.uint16)])
def test_noalloc_equals_workspace_call_twice(params, k, dtype):
    from bitlev import hip
    assert hip.available(), "no HIP device"
    big = _sets(1, 300, 5000, (0, 200), (0, 220))
    small = _sets(2, 37, 900, (1, 70), (0, 90))
    for pq, pc in (big, small):
        print(f"noalloc nq={len(pq[1]) - 1} nc={len(pc[1]) - 1} k={k} params={params}", flush=True)
        want = _run(hip, "tuned", pq, pc, k, dtype, params)
        got = _run(hip, "noalloc", pq, pc, k, dtype, params)
        np.testing.assert_array_equal(got, want)
    # and the second (smaller) one against the oracle
    pq, pc = small
    np.testing.assert_array_equal(got, O.saturate(O.cross(pq, pc, k=k), np.dtype(dtype)))


def test_noalloc_refuses_more_than_its_scratch():
    from bitlev import _capi, _edpy, hip
    assert hip.available(), "no HIP device"
    pq, pc = _sets(3, 2, 10, (5, 9), (5, 9))
    Q, C = hip.Packed(*pq), hip.Packed(*pc)
    out = hip.DeviceArray.empty((Q.n, C.n), np.int32)
    # 600 queries of up to 512 bytes need 600 x 16 KiB: above the 8 MiB array
    need = _edpy.bl_cdist_workspace_size(600, 512)
    assert need > 8 << 20
    rc = _edpy.bl_cdist_noalloc(Q.d_buf.ptr, Q.d_off.ptr, 600, C.stripes().desc, 512, out.ptr, 0,
                                -1, None, None, None)
    assert rc == _capi.BL_EINVAL
    assert b"static scratch" in _edpy.bl_last_error()
