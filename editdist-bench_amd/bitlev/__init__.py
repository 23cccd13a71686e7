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
The top-level distance / pairwise / cdist forward to `hip`.
"""
from __future__ import annotations

import numpy as np

__version__ = "0.1.0"
__all__ = ["distance", "pairwise", "cdist", "hip", "gpu", "available"]

from . import hip  # noqa: E402
from . import hip as gpu  # noqa: E402

available = hip.available


def distance(a, b, max_dist=None) -> int:
    return hip.distance(a, b, max_dist=max_dist)


def _method(method):
    if method not in ("auto", "hip", "gpu"):
        raise ValueError("method must be 'auto' or 'hip' (this package has no CPU engine)")


def pairwise(a, b, max_dist=None, dtype=np.int32, method: str = "auto", return_device: bool = False):
    _method(method)
    return hip.pairwise(a, b, max_dist=max_dist, dtype=dtype, return_device=return_device)


def cdist(queries, choices, max_dist=None, dtype=np.int32, method: str = "auto",
          return_device: bool = False):
    _method(method)
    return hip.cdist(queries, choices, max_dist=max_dist, dtype=dtype, return_device=return_device)
