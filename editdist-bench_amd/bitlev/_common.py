"""Argument checks shared by the backend (same rules as bitlev/_common.py:9-33
of the reference: max_dist None -> -1, negative -> ValueError, clamp to 2**30;
dtype one of int32 / uint16 / uint8), and the saturating conversion of the
CPU methods' int32 results."""
from __future__ import annotations

import numpy as np

DTYPES = (np.int32, np.uint16, np.uint8)
DTYPE_CODE = {np.dtype(np.int32): 0, np.dtype(np.uint16): 1, np.dtype(np.uint8): 2}


def check_k(max_dist) -> int:
    if max_dist is None:
        return -1
    k = int(max_dist)
    if k < 0:
        raise ValueError("max_dist must be >= 0 (or None)")
    return min(k, 2**30)


def check_dtype(dtype) -> np.dtype:
    dt = np.dtype(dtype)
    if dt not in DTYPE_CODE:
        raise ValueError("dtype must be numpy int32, uint16 or uint8")
    return dt


def as_list(x):
    return x if isinstance(x, list) else list(x)


def saturate(out: np.ndarray, dt) -> np.ndarray:
    """int32 results in the requested dtype, narrower types saturating at
    their maximum (bitlev/_common.py:25-30)."""
    dt = np.dtype(dt)
    if dt == np.int32:
        return out
    return np.minimum(out, np.iinfo(dt).max).astype(dt)
