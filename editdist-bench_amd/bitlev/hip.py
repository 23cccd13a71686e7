"""bitlev.hip -- exact edit distance on AMD Instinct MI355X (gfx950).

Drop-in for the reference's `bitlev.gpu` (bitlev/gpu.py:71-138): same
functions, arguments, results, dtypes and exceptions, with hand-written HIP
kernels behind the C ABI of libbitlev_hip.so instead of CuPy/NVRTC.

    from bitlev import hip
    d = hip.cdist(queries, choices, max_dist=2, dtype="uint8")
    d = hip.pairwise(a, b, return_device=True)        # DeviceArray, .get() -> numpy

Every distance is computed on the GPU, including the strings the reference
hands to its CPU core (pairs whose shorter string exceeds 256 bytes and cdist
queries longer than 512 bytes): those run on the wavefront kernel. There is no
CPU fallback; without the built library or a GPU the calls raise.

Two forms compute the same results:
  * "batched" (default): one call per batch, no host round trip -- on-device
    length bucketing and one persistent launch (bl_pairs / bl_cdist);
  * "plain": the host splits pairs and queries into the reference's length
    classes (bitlev/gpu.py:85-86, :116-118) and launches one kernel per class.
Select with the module attribute FORM or the `form` argument of the *_into
functions.
"""
from __future__ import annotations

import numpy as np

from . import _edpy
import ctypes

from ._capi import BackendUnavailable, CdistParams, PairsParams, Stripes, call, lib
from ._common import DTYPE_CODE, as_list, check_dtype, check_k
from .device import DeviceArray, DeviceBuffer

PAIR_CAPS = (16, 32, 64, 128, 256)   # pairs kernels: shorter string <= CAP
CROSS_NWS = (1, 2, 4, 8)             # cdist kernels: query <= 64 * NW
BATCHED_MAX_QUERY = 512              # longer cdist queries take the wavefront kernel
SORT_BLOCK = 4096                    # length-ordered stripes: strings sorted within blocks of this many
FORM = "batched"                     # default form of pairwise() / cdist()

__all__ = ["available", "distance", "pairwise", "cdist", "Packed", "Plan", "pairs_plan",
           "cdist_plan", "pairs_into", "cdist_into", "pairs_batched_into", "cdist_batched_into",
           "run_pairs", "run_cdist", "Workspace", "Stream", "Graph", "pair_classes", "query_classes",
           "ChoiceStripes", "cdist_matches", "cdist_matches_into",
           "DeviceArray", "BackendUnavailable", "FORM"]


def available() -> bool:
    """True when the HIP library loads and a GPU is visible. Never raises
    (bitlev/gpu.py:36-41)."""
    try:
        return lib().bl_device_count() > 0
    except Exception:
        return False


def _require_device():
    if lib().bl_device_count() <= 0:
        raise BackendUnavailable("no HIP device visible")


# ------------------------------------------------------------------ host side --
def pack(strings) -> tuple[np.ndarray, np.ndarray]:
    """str (chars <= U+00FF) / bytes -> (uint8 buffer, int64 offsets[n + 1]).
    Long lists are packed by a thread pool (bitlev.ingest.pack)."""
    if isinstance(strings, (list, tuple)) and len(strings) >= _PARALLEL_PACK:
        from . import ingest
        return ingest.pack(strings)
    buf, off = _edpy.pack(strings)
    return np.frombuffer(buf, dtype=np.uint8), np.frombuffer(off, dtype=np.int64)


_PARALLEL_PACK = 1 << 15


def pair_classes(off_a: np.ndarray, off_b: np.ndarray):
    """Pair indices per pairs kernel, by the shorter length before trimming,
    and the rest (shorter > 256) for the long kernel (bench/cudamyers.py:23-28)."""
    m = np.minimum(np.diff(off_a), np.diff(off_b))
    cls = np.searchsorted(np.array(PAIR_CAPS), m)
    groups = [np.nonzero(cls == c)[0].astype(np.int32) for c in range(len(PAIR_CAPS))]
    return groups, np.nonzero(cls == len(PAIR_CAPS))[0].astype(np.int32)


def query_classes(off_q: np.ndarray):
    """Query indices per cdist kernel by words = ceil(len / 64) (an empty query
    goes to NW 1), and the rest (> 512) for the long kernel (bench/cudamyers.py:31-36)."""
    words = (np.diff(off_q) + 63) // 64
    cls = np.searchsorted(np.array(CROSS_NWS), words)
    groups = [np.nonzero(cls == c)[0].astype(np.int32) for c in range(len(CROSS_NWS))]
    return groups, np.nonzero(cls == len(CROSS_NWS))[0].astype(np.int32)


class Packed:
    """A packed string set resident on the device (buffer + int64 offsets),
    with its host offsets kept for the class split. Build once, reuse across
    calls (the reference's CrossState/PairsState, bench/cudamyers.py:83-151)."""

    def __init__(self, buf: np.ndarray, off: np.ndarray, stream=None):
        self.h_off = np.ascontiguousarray(off, dtype=np.int64)
        self.n = len(self.h_off) - 1
        self.lengths = np.diff(self.h_off)
        self.max_len = int(self.lengths.max()) if self.n else 0
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        if buf.size == 0:
            buf = np.zeros(1, np.uint8)
        self.nbytes = int(buf.nbytes)
        self.d_buf = DeviceBuffer.from_host(buf, stream)
        self.d_off = DeviceBuffer.from_host(self.h_off, stream)

        self._stripes = None
        self._sorted = None

    @classmethod
    def from_strings(cls, strings, stream=None) -> "Packed":
        return cls(*pack(strings), stream=stream)

    @classmethod
    def from_parts(cls, parts, stream=None) -> "Packed":
        """One resident set of several packed host sets (buf, off) in turn,
        each part's bytes copied straight to its place in the device buffer
        (no host concatenation: a fresh 50 MB host array costs more in page
        faults than its upload). The parts must stay alive and unchanged
        until `stream` has run the copies."""
        self = cls.__new__(cls)
        offs, base = [np.zeros(1, np.int64)], 0
        spans = []
        for buf, off in parts:
            off = np.asarray(off, np.int64)
            n = int(off[-1] - off[0])
            spans.append((buf, int(off[0]), n, base))
            offs.append(off[1:] - off[0] + base)
            base += n
        self.h_off = np.concatenate(offs)
        self.n = len(self.h_off) - 1
        self.lengths = np.diff(self.h_off)
        self.max_len = int(self.lengths.max()) if self.n else 0
        self.nbytes = max(base, 1)
        self.d_buf = DeviceBuffer(self.nbytes)
        for buf, o, n, at in spans:
            if n:
                self.d_buf.upload(np.asarray(buf, np.uint8)[o:o + n], stream, at)
        self.d_off = DeviceBuffer.from_host(self.h_off, stream)
        self._stripes = None
        self._sorted = None
        return self

    def stripes(self, stream=None, order: str = "set", grams3: bool = False) -> "ChoiceStripes":
        """The set as search-kernel stripes (built on the device on first use,
        on `stream`, which is then synchronised; kept with the set).
        order="set": slot c holds string c, with the byte-class and 2-gram
        profiles (every search mode; grams3 adds the 3-gram profiles, and
        rebuilds a layout made without them); order="length": sorted by
        length within blocks of SORT_BLOCK strings, no profiles (exact scans
        only)."""
        if order == "length":
            if self._sorted is None:
                if self.n < 2 or bool(np.all(self.lengths[1:] >= self.lengths[:-1])):
                    # already in length order (a serving set): the set's own stripes
                    self._sorted = self.stripes(stream)
                else:
                    self._sorted = ChoiceStripes(self, stream, profiles=False, sort_block=SORT_BLOCK)
            return self._sorted
        if order != "set":
            raise ValueError("order must be 'set' or 'length'")
        if self._stripes is None or (grams3 and self._stripes.d_counts3 is None):
            self._stripes = ChoiceStripes(self, stream, grams3=grams3)
        return self._stripes

    @property
    def device_bytes(self) -> int:
        b = self.d_buf.nbytes + self.d_off.nbytes
        for st in {id(x): x for x in (self._stripes, self._sorted) if x is not None}.values():
            b += st.nbytes
        return b


class ChoiceStripes:
    """A packed set interleaved by dword in groups of 64 strings (bl_stripes):
    the layout the cdist kernels read, so that a wave's lanes load 256
    contiguous bytes per step, plus each string's 16-byte byte-class and
    2-gram profiles (the batched searches' lower bounds; profiles=False
    leaves them out) and, with grams3, the 64-byte 3-gram profiles (a third
    bound for small alphabets; off by default: on the 256 x 4.2M DNA search
    at k = 8 it settles what the 2-gram bound leaves but its per-pair profile
    loads cost as much as the scans they save, 1.83 -> 1.81 ms, and the
    set's build takes 14 ms instead of a few). sort_block > 0: the strings in
    length order within blocks of that many (exact scans)."""

    def __init__(self, P: Packed, stream=None, profiles: bool = True, sort_block: int = 0,
                 grams3: bool = False):
        n = P.n
        groups = (n + 63) // 64
        self.d_order = None
        if sort_block:
            gbase = np.empty(groups + 1, np.int64)
            order = np.empty(max(n, 1), np.int32)
            total = call("bl_stripes_layout_sorted", P.h_off.ctypes.data, n, int(sort_block),
                         order.ctypes.data, gbase.ctypes.data)
            self.d_order = DeviceBuffer.from_host(order, stream)
            self.d_gbase = DeviceBuffer.from_host(gbase, stream)
        else:
            # the group bases from the resident offsets (synchronises the stream)
            self.d_gbase = DeviceBuffer(8 * (groups + 1))
            work = DeviceBuffer(call("bl_stripes_layout_workspace_size", n))
            total = call("bl_stripes_layout_device", P.d_off.ptr, n, self.d_gbase.ptr, work.ptr,
                         work.nbytes, stream)
            del work
        self.d_words = DeviceBuffer(max(int(total), 1) * 4)
        self.d_len = DeviceBuffer(max(n, 1) * 4)
        self.d_counts = DeviceBuffer(max(n, 1) * 16) if profiles else None
        self.d_counts2 = DeviceBuffer(max(n, 1) * 16) if profiles else None
        self.d_counts3 = DeviceBuffer(max(n, 1) * 64) if profiles and grams3 else None
        counts = self.d_counts.ptr if profiles else None
        counts2 = self.d_counts2.ptr if profiles else None
        counts3 = self.d_counts3.ptr if self.d_counts3 is not None else None
        call("bl_stripes_build_order", P.d_buf.ptr, P.d_off.ptr, n,
             self.d_order.ptr if sort_block else None, self.d_gbase.ptr, self.d_words.ptr,
             self.d_len.ptr, counts, counts2, counts3, stream)
        # built once per set; finish here so that any stream may read them next
        call("bl_stream_sync", stream)
        self.nbytes = self.d_gbase.nbytes + self.d_words.nbytes + self.d_len.nbytes
        if profiles:
            self.nbytes += self.d_counts.nbytes + self.d_counts2.nbytes
        if self.d_counts3 is not None:
            self.nbytes += self.d_counts3.nbytes
        if sort_block:
            self.nbytes += self.d_order.nbytes
        self.desc = Stripes(self.d_words.ptr, self.d_gbase.ptr, self.d_len.ptr, n, counts, counts2,
                            self.d_order.ptr if sort_block else None, counts3, P.d_buf.ptr, P.d_off.ptr)
        self._packed = P            # the buffers desc.buf / desc.off point into

    @property
    def ref(self):
        return self.desc


def _idx_buffers(groups, stream=None):
    return [DeviceBuffer.from_host(g, stream) if len(g) else None for g in groups]


class Plan:
    """Class split of one call shape, with its index lists on the device.
    Built once per (A, B) pair set or per query set and reused by every call
    on the same resident inputs, so a repeated call is kernel launches only.
    The index lists are uploaded on `stream` (the stream the calls run on);
    without one the upload has finished when the plan is returned."""

    def __init__(self, groups, rest, max_len, stream=None):
        self.groups = groups
        self.bufs = _idx_buffers(groups, stream)
        self.rest = rest
        self.rest_buf = DeviceBuffer.from_host(rest, stream) if len(rest) else None
        self.max_len = max_len


def pairs_plan(A: Packed, B: Packed, stream=None) -> Plan:
    groups, rest = pair_classes(A.h_off, B.h_off)
    mx = int(max(A.lengths[rest].max(), B.lengths[rest].max())) if len(rest) else 0
    return Plan(groups, rest, mx, stream)


def cdist_plan(Q: Packed, C: Packed, stream=None) -> Plan:
    groups, rest = query_classes(Q.h_off)
    mx = int(max(Q.lengths[rest].max(), C.max_len)) if len(rest) else 0
    return Plan(groups, rest, mx, stream)


def pairs_into(A: Packed, B: Packed, out: DeviceArray, k: int = -1, trim: bool = True,
               stream=None, plan: Plan | None = None) -> Plan:
    """Launch the plain pairs kernels: out[i] = d(A[i], B[i]) (min(d, k+1) for
    k >= 0) in out's dtype. Asynchronous on `stream`; keep the returned plan
    alive until the work is done (it owns the index lists)."""
    if A.n != B.n:
        raise ValueError(f"length mismatch: {A.n} vs {B.n}")
    if plan is None:
        plan = pairs_plan(A, B, stream)
    if A.n == 0:
        return plan
    dt = DTYPE_CODE[out.dtype]
    for cap, g, b in zip(PAIR_CAPS, plan.groups, plan.bufs):
        if b is not None:
            call("bl_pairs_class", cap, A.d_buf.ptr, A.d_off.ptr, B.d_buf.ptr, B.d_off.ptr, b.ptr,
                 len(g), out.ptr, dt, int(trim), int(k), stream)
    if plan.rest_buf is not None:
        call("bl_long", 1, A.d_buf.ptr, A.d_off.ptr, B.d_buf.ptr, B.d_off.ptr, plan.rest_buf.ptr,
             len(plan.rest), 0, out.ptr, dt, int(trim), int(k), plan.max_len, stream)
    return plan


def cdist_into(Q: Packed, C: Packed, out: DeviceArray, k: int = -1, stream=None,
               plan: Plan | None = None, stripes: bool = False) -> Plan:
    """Launch the plain cdist kernels into the row-major (Q.n, C.n) `out`:
    one bl_cdist_class launch per query class over the packed choices, the
    reference's launches one for one (bitlev/gpu.py:121-130); stripes=True
    reads the choice set's resident stripes instead (bl_cdist_class_stripes)."""
    if plan is None:
        plan = cdist_plan(Q, C, stream)
    if Q.n == 0 or C.n == 0:
        return plan
    dt = DTYPE_CODE[out.dtype]
    st = C.stripes(stream) if stripes else None
    for nw, g, b in zip(CROSS_NWS, plan.groups, plan.bufs):
        if b is None:
            continue
        if st is not None:
            call("bl_cdist_class_stripes", nw, Q.d_buf.ptr, Q.d_off.ptr, b.ptr, len(g), st.ref,
                 out.ptr, dt, int(k), stream)
        else:
            call("bl_cdist_class", nw, Q.d_buf.ptr, Q.d_off.ptr, b.ptr, len(g), C.d_buf.ptr,
                 C.d_off.ptr, C.n, out.ptr, dt, int(k), stream)
    if plan.rest_buf is not None:
        call("bl_long", 0, Q.d_buf.ptr, Q.d_off.ptr, C.d_buf.ptr, C.d_off.ptr, plan.rest_buf.ptr,
             len(plan.rest), C.n, out.ptr, dt, 0, int(k), plan.max_len, stream)
    return plan


class Workspace:
    """Device workspace of the batched entry points, grown on demand and
    reused across calls (calls sharing one workspace must be ordered on one
    stream)."""

    def __init__(self):
        self.buf = None

    def get(self, nbytes: int) -> DeviceBuffer:
        if self.buf is None or self.buf.nbytes < nbytes:
            self.buf = DeviceBuffer(max(int(nbytes), 256))
        return self.buf


_default_ws = {}


def _workspace(stream):
    ws = _default_ws.get(stream)
    if ws is None:
        ws = _default_ws[stream] = Workspace()
    return ws


def _cdist_params(params):
    if params is None:
        return None
    p = CdistParams()
    p.waves_per_cu = int(params.get("waves_per_cu", 0))
    p.tile = int(params.get("tile", 0))
    p.reserved[0] = {"strided": 1, "swar": 2, "gap": 3, "count": 4, "exact": 5,
                     "queue": 6}.get(params.get("mode"), 0)
    p.reserved[2] = int(params.get("qgroup", 0))
    p.reserved[1] = int(params.get("grab", 0))
    p.reserved[3] = ((0 if params.get("gram", True) else 1) | (0 if params.get("static", True) else 2)
                     | (0 if params.get("list", True) else 4))
    return p


def _pairs_params(params):
    if params is None:
        return None
    p = PairsParams()
    p.waves_per_cu = int(params.get("waves_per_cu", 0))
    p.block = int(params.get("block", 0))
    p.bin_width = int(params.get("bin_width", 0))
    p.reserved[0] = int(params.get("chunk", 0))
    p.reserved[1] = 1 if params.get("mode") == "direct" else 0
    p.reserved[2] = int(params.get("grab", 0))
    p.reserved[3] = (int(params.get("band", 0)) | (int(params.get("long_units", 0)) << 4)
                     | ((1 if params.get("unit64", False) else 0) << 8))
    return p


def pairs_batched_into(A: Packed, B: Packed, out: DeviceArray, k: int = -1, trim: bool = True,
                       stream=None, workspace: Workspace | None = None, params=None,
                       kvec: DeviceBuffer | None = None):
    """One bl_pairs pipeline (bucketing + persistent launch) into `out`;
    kvec: optional device int32 per-pair thresholds (overrides k).
    Trimming never changes a distance (d(xPy, xQy) = d(P, Q)), so where the
    caller allows it (trim=True) the launch shape may still switch it off
    (params["trim"] = False) when scanning the affixes is cheaper than
    comparing them."""
    if A.n != B.n:
        raise ValueError(f"length mismatch: {A.n} vs {B.n}")
    if A.n == 0:
        return
    if params is not None and not params.get("trim", True):
        trim = False
    max_len = max(A.max_len, B.max_len)
    ws = (workspace or _workspace(stream)).get(lib().bl_pairs_workspace_size(A.n, max_len))
    call("bl_pairs_tuned", A.d_buf.ptr, A.d_off.ptr, B.d_buf.ptr, B.d_off.ptr, A.n, max_len,
         out.ptr, DTYPE_CODE[out.dtype], int(trim), int(k), kvec.ptr if kvec else None, ws.ptr,
         ws.nbytes, _pairs_params(params), stream)


def cdist_batched_into(Q: Packed, C: Packed, out: DeviceArray, k: int = -1, stream=None,
                       workspace: Workspace | None = None, params=None, plan: Plan | None = None,
                       kvec: DeviceBuffer | None = None):
    """One bl_cdist persistent launch into the row-major (Q.n, C.n) `out`;
    queries above 512 bytes follow on the wavefront kernel (bl_long).
    kvec: optional device int32 per-query thresholds (overrides k; queries
    above 512 bytes then need a single k)."""
    if Q.n == 0 or C.n == 0:
        return plan
    mq = min(Q.max_len, BATCHED_MAX_QUERY)
    ws = (workspace or _workspace(stream)).get(lib().bl_cdist_workspace_size(Q.n, mq))
    # exact scans (the automatic mode of an exact search, or asked for) read
    # the set in length order, so that the lanes of a wave finish together
    mode = params.get("mode") if params else None
    exact = mode == "exact" or (mode is None and k < 0 and kvec is None)
    st = C.stripes(stream, "length" if exact else "set")
    call("bl_cdist_tuned", Q.d_buf.ptr, Q.d_off.ptr, Q.n, st.ref, mq, out.ptr,
         DTYPE_CODE[out.dtype], int(k), kvec.ptr if kvec else None, ws.ptr, ws.nbytes,
         _cdist_params(params), stream)
    if Q.max_len > BATCHED_MAX_QUERY:
        if kvec is not None:
            raise ValueError("per-query thresholds need every query <= 512 bytes")
        if plan is None:
            plan = cdist_plan(Q, C, stream)
        call("bl_long", 0, Q.d_buf.ptr, Q.d_off.ptr, C.d_buf.ptr, C.d_off.ptr, plan.rest_buf.ptr,
             len(plan.rest), C.n, out.ptr, DTYPE_CODE[out.dtype], 0, int(k), plan.max_len, stream)
    return plan


def default_pairs_params(k: int):
    """Launch shape of an untuned batched pairwise call: with a threshold a
    64-row band window always reaches (0 <= k <= 30), band tiles from two
    32-row units with text-length bins of 8 (the 16.8M uneven pairs at k = 8:
    9.0 -> 4.7 ms; at k = 3: 4.2 -> 3.0 ms); otherwise the library's
    defaults."""
    if 0 <= k <= 30:
        return {"band": 2, "bin_width": 8, "chunk": 16384}
    return None


def run_pairs(A: Packed, B: Packed, out: DeviceArray, k: int = -1, trim: bool = True,
              stream=None, form: str | None = None, plan=None, workspace=None, params=None):
    """pairs in the chosen form; returns what must stay alive until the work is done."""
    form = form or FORM
    if form == "batched":
        if params is None:
            params = default_pairs_params(k)
        pairs_batched_into(A, B, out, k, trim, stream, workspace, params)
        return None
    if form == "plain":
        return pairs_into(A, B, out, k, trim, stream, plan)
    raise ValueError("form must be 'batched' or 'plain'")


def default_cdist_params(nq: int, nc: int, k: int, cus: int = 256):
    """Launch shape of an untuned batched search: with a threshold, the count
    kernel (query groups of up to 64; blocks sized so the 16 waves per CU get
    about two tiles each -- the shape the autotuner picks for the
    1,024 x 1e6 and 256 x 4.2e6 searches); exact searches keep the library's
    defaults."""
    if k < 0 or nq == 0 or nc == 0:
        return None
    g = min(64, nq)
    waves = 16 * cus
    tile = nc * ((nq + g - 1) // g) // (2 * waves)
    tile = min(8192, max(1024, tile // 256 * 256))
    return {"mode": "count", "waves_per_cu": 16, "tile": tile, "qgroup": g}


def run_cdist(Q: Packed, C: Packed, out: DeviceArray, k: int = -1, stream=None,
              form: str | None = None, plan=None, workspace=None, params=None):
    form = form or FORM
    if form == "batched":
        if params is None and C.stripes(stream).desc.counts:
            params = default_cdist_params(Q.n, C.n, k)
        return cdist_batched_into(Q, C, out, k, stream, workspace, params, plan)
    if form == "plain":
        return cdist_into(Q, C, out, k, stream, plan)
    raise ValueError("form must be 'batched' or 'plain'")


class Stream:
    """A HIP stream (default priority) owned by Python."""

    def __init__(self):
        p = ctypes.c_void_p()
        call("bl_stream_create", p)
        self.ptr = p.value

    def sync(self):
        call("bl_stream_sync", self.ptr)

    def __del__(self):
        try:
            if self.ptr:
                lib().bl_stream_destroy(self.ptr)
        except Exception:
            pass


class Graph:
    """Work issued by fn() on `stream`, captured into a HIP graph and replayed
    with launch(). fn runs once eagerly first, so every buffer it needs (the
    workspace included) exists before capture; the captured calls must then
    allocate nothing and sync nothing."""

    def __init__(self, fn, stream: Stream, warm: bool = True):
        self.stream = stream
        self.fn = fn          # holds what the captured calls read and write (workspace, outputs)
        if warm:
            fn()
            stream.sync()
        call("bl_graph_begin", stream.ptr)
        ge = ctypes.c_void_p()
        try:
            fn()
        finally:
            rc = lib().bl_graph_end(stream.ptr, ge)
        from ._capi import check
        check(rc, "bl_graph_end")
        self.exec = ge.value

    def launch(self):
        call("bl_graph_launch", self.exec, self.stream.ptr)

    def __del__(self):
        try:
            if self.exec:
                lib().bl_graph_destroy(self.exec)
        except Exception:
            pass


def _finish(out: DeviceArray, return_device: bool, stream=None):
    call("bl_stream_sync", stream)
    return out if return_device else out.get(stream)


# ------------------------------------------------------------------ public API --
# pairwise calls of at least this many pairs (results to the host) go through
# the staged ingest pipeline (bitlev.ingest); smaller ones pack in one piece
STAGED_PAIRS = 1 << 16

def pairwise(a, b, max_dist=None, dtype=np.int32, return_device: bool = False):
    """out[i] = d(a[i], b[i]) on the GPU (bitlev/gpu.py:71-103).
    return_device=True returns a DeviceArray (`.get()` gives numpy)."""
    a, b = as_list(a), as_list(b)
    if len(a) != len(b):
        raise ValueError(f"length mismatch: {len(a)} vs {len(b)}")
    k, dt = check_k(max_dist), check_dtype(dtype)
    n = len(a)
    if n >= STAGED_PAIRS and not return_device:
        _require_device()
        from . import ingest                   # packing, copies and kernels overlapped
        return ingest.pairwise(a, b, k, dt, True)
    pa, pb = pack(a), pack(b)
    _require_device()
    out = DeviceArray.empty((n,), dt)
    if n == 0:
        return out if return_device else np.empty(0, dt)
    A, B = Packed(*pa), Packed(*pb)
    keep = run_pairs(A, B, out, k, True, None)
    res = _finish(out, return_device)
    del keep
    return res


def cdist(queries, choices, max_dist=None, dtype=np.int32, return_device: bool = False):
    """out[i, j] = d(queries[i], choices[j]) on the GPU (bitlev/gpu.py:106-138)."""
    q, c = as_list(queries), as_list(choices)
    k, dt = check_k(max_dist), check_dtype(dtype)
    nq, nc = len(q), len(c)
    # a long choice list is packed straight into pinned memory and uploaded
    # from there (bitlev.ingest); a short one the ordinary way
    staged = nc >= _PARALLEL_PACK
    pq, pc = pack(q), (None if staged else pack(c))
    _require_device()
    out = DeviceArray.empty((nq, nc), dt)
    if nq == 0 or nc == 0:
        return out if return_device else np.empty((nq, nc), dt)
    Q = Packed(*pq)
    if staged:
        from . import ingest
        C = ingest.upload_packed(c)
    else:
        C = Packed(*pc)
    keep = run_cdist(Q, C, out, k, None)
    res = _finish(out, return_device)
    del keep
    return res


def cdist_matches_into(Q: Packed, C: Packed, k: int, stream=None,
                       workspace: Workspace | None = None, params=None,
                       kvec: DeviceBuffer | None = None, capacity: int | None = None):
    """Every (q, c) with d(Q[q], C[c]) <= k as three device arrays (q int32,
    c int64, d int32) in no particular order, through bl_cdist_matches: the
    search without the dense Q.n x C.n result. Grows the buffers and runs
    again when the first capacity was too small; synchronises `stream`."""
    if k < 0 and kvec is None:
        raise ValueError("a match list needs max_dist >= 0")
    mq = min(Q.max_len, BATCHED_MAX_QUERY)
    if Q.max_len > BATCHED_MAX_QUERY:
        raise ValueError("match lists take queries of at most 512 bytes")
    ws = (workspace or _workspace(stream)).get(lib().bl_cdist_workspace_size(Q.n, mq))
    cap = int(capacity if capacity is not None else min(max(Q.n * C.n, 1), 1 << 20))
    count = DeviceBuffer(8)
    host_count = np.zeros(1, np.uint64)
    while True:
        dq = DeviceArray.empty((max(cap, 1),), np.int32)
        dc = DeviceArray.empty((max(cap, 1),), np.int64)
        dd = DeviceArray.empty((max(cap, 1),), np.int32)
        call("bl_cdist_matches", Q.d_buf.ptr, Q.d_off.ptr, Q.n, C.stripes(stream).ref, mq, int(k),
             kvec.ptr if kvec else None, dq.ptr, dc.ptr, dd.ptr, cap, count.ptr, ws.ptr, ws.nbytes,
             _cdist_params(params), stream)
        call("bl_memcpy_dtoh", host_count.ctypes.data, count.ptr, 8, stream)
        call("bl_stream_sync", stream)
        n = int(host_count[0])
        if n <= cap:
            return DeviceArray((n,), np.int32, dq._buf), DeviceArray((n,), np.int64, dc._buf), \
                DeviceArray((n,), np.int32, dd._buf)
        cap = n


def cdist_matches(queries, choices, max_dist, return_device: bool = False):
    """Every pair within max_dist: (q, c, d) arrays with d(queries[q],
    choices[c]) = d <= max_dist, sorted by (q, c) when returned to the host
    (device arrays come in no particular order). The reference's cdist with
    max_dist gives the same information as a dense matrix (bitlev/gpu.py:106-138)."""
    q, c = as_list(queries), as_list(choices)
    k = check_k(max_dist)
    if k < 0:
        raise ValueError("a match list needs max_dist >= 0")
    _require_device()
    if len(q) == 0 or len(c) == 0:
        z = (np.empty(0, np.int32), np.empty(0, np.int64), np.empty(0, np.int32))
        return z
    Q, C = Packed(*pack(q)), Packed(*pack(c))
    dq, dc, dd = cdist_matches_into(Q, C, k)
    if return_device:
        return dq, dc, dd
    hq, hc, hd = dq.get(), dc.get(), dd.get()
    order = np.lexsort((hc, hq))
    return hq[order], hc[order], hd[order]


def distance(a, b, max_dist=None) -> int:
    """Edit distance of one pair (min(d, max_dist + 1) with max_dist),
    bitlev/__init__.py:41-45."""
    return int(pairwise([a], [b], max_dist=max_dist)[0])
