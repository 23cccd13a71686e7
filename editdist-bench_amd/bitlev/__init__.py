"""bitlev (MI355X backend): exact unit-cost Levenshtein distance for batches of
short byte strings, computed by hand-written HIP kernels for AMD gfx950.

    from bitlev import hip            # also importable as bitlev.gpu
    hip.distance("kitten", "sitting")                       # 3
    hip.pairwise(["abc", "kitten"], ["abd", "sitting"])     # array([1, 3])
    hip.cdist(queries, choices, max_dist=2, dtype="uint8")  # len(queries) x len(choices)

Strings are compared byte by byte: `str` whose characters are all <= U+00FF,
or `bytes`. `max_dist=k` returns min(distance, k + 1); `dtype` is int32,
uint16 or uint8, narrower types saturating at their maximum. Same contract as
the reference's bitlev (bitlev/__init__.py:1-21, bitlev/gpu.py).

Submodules: `hip` (the drop-in per-call API), `gpu` (alias of `hip`, so code
written against the reference's `bitlev.gpu` runs unchanged).
The top-level distance / pairwise / cdist forward to `hip`; their CPU
methods ("scalar", "simd": AVX2) run the CPU engine of `_edpy` only when a
caller names them.
"""
from __future__ import annotations

import numpy as np

__version__ = "0.1.0"
__all__ = ["distance", "pairwise", "cdist", "hip", "gpu", "available", "simd_available"]

from . import _edpy  # noqa: E402
from . import hip  # noqa: E402
from . import hip as gpu  # noqa: E402
from ._common import as_list, check_dtype, check_k, saturate  # noqa: E402

available = hip.available


def simd_available() -> bool:
    """True when the CPU engine's AVX2 kernels can run on this host (the
    reference's NEON kernels, bitlev/__init__.py:33-35 there)."""
    return bool(_edpy.simd_available())


def distance(a, b, max_dist=None) -> int:
    return hip.distance(a, b, max_dist=max_dist)


_GPU_METHODS = ("auto", "hip", "gpu")


def _method(method):
    if method not in _GPU_METHODS + ("scalar", "simd"):
        raise ValueError("method must be 'auto' (GPU), 'hip', 'scalar' or 'simd'")
    if method == "simd" and not simd_available():
        raise ValueError("method='simd' needs AVX2")
    return method


def pairwise(a, b, max_dist=None, dtype=np.int32, method: str = "auto", return_device: bool = False):
    """out[i] = d(a[i], b[i]). method "auto" (or "hip"/"gpu") runs on the GPU;
    "scalar" and "simd" run the CPU engine the reference's CPU methods name
    (bitlev/__init__.py:53-77 there): one pair at a time, or AVX2 lanes."""
    m = _method(method)
    if m in _GPU_METHODS:
        return hip.pairwise(a, b, max_dist=max_dist, dtype=dtype, return_device=return_device)
    a, b = as_list(a), as_list(b)
    if len(a) != len(b):
        raise ValueError(f"length mismatch: {len(a)} vs {len(b)}")
    k, dt = check_k(max_dist), check_dtype(dtype)
    out = np.empty(len(a), dtype=np.int32)
    if m == "simd":
        _edpy.pairs_simd(a, b, out, True, 4)
        if k >= 0:
            np.minimum(out, k + 1, out=out)
    else:
        _edpy.pairs(a, b, out, True, k)
    return saturate(out, dt)


def cdist(queries, choices, max_dist=None, dtype=np.int32, method: str = "auto",
          return_device: bool = False):
    """out[i, j] = d(queries[i], choices[j]); methods as pairwise
    (bitlev/__init__.py:80-101 there)."""
    m = _method(method)
    if m in _GPU_METHODS:
        return hip.cdist(queries, choices, max_dist=max_dist, dtype=dtype, return_device=return_device)
    q, c = as_list(queries), as_list(choices)
    k, dt = check_k(max_dist), check_dtype(dtype)
    out = np.empty(len(q) * len(c), dtype=np.int32)
    if m == "simd":
        _edpy.cross_simd(q, c, out, 4, k)
    else:
        _edpy.cross(q, c, out, k)
    return saturate(out.reshape(len(q), len(c)), dt)
