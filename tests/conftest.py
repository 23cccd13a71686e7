"""Shared test setup: import paths, the `gpu` marker and golden-vector access."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "editdist-bench_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_pair_sets(g):
    names = sorted({k.split("/")[1] for k in g if k.startswith("pairs/")})
    return names


def golden_cross_sets(g):
    return sorted({k.split("/")[1] for k in g if k.startswith("cross/")})


def pairs_of(g, name):
    return ((g[f"pairs/{name}/buf_a"], g[f"pairs/{name}/off_a"]),
            (g[f"pairs/{name}/buf_b"], g[f"pairs/{name}/off_b"]))


def cross_of(g, name):
    return ((g[f"cross/{name}/buf_q"], g[f"cross/{name}/off_q"]),
            (g[f"cross/{name}/buf_c"], g[f"cross/{name}/off_c"]))


def unpack(buf, off):
    return [bytes(buf[off[i]:off[i + 1]]) for i in range(len(off) - 1)]
