"""The persistent kernels' work-queue schedule (queue_of / queue_run in
csrc/bl_device.h, built for the host) played out on the CPU: waves of one
launch draw runs of tiles in random interleavings, each wave from its own
queue, each draw taking the queue counter's value and adding one, a wave
stopping at its first tile past the end -- as Tickets does on the device.
Every tile must be handed out exactly once, for launches of fewer waves than
queues too. A planted schedule that uses all 16 queues with fewer waves leaves
tiles undone, which the same check catches. No GPU needed."""
from __future__ import annotations

import ctypes
import os
import random

import pytest

from conftest import PKG

PATH = os.path.join(PKG, "bitlev", "libbl_hostcheck.so")
QUEUES = 16


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(PATH):
        pytest.fail(f"{PATH} not built (make -C editdist-bench_amd)")
    L = ctypes.CDLL(PATH)
    L.blh_queue_of.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.blh_queue_of.restype = ctypes.c_int
    L.blh_queue_run.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int64]
    L.blh_queue_run.restype = ctypes.c_int64
    return L


def play(queue_of, queue_run, waves, grab, ntiles, first_static, seed):
    """Tiles handed out by a launch of `waves` waves, in a random order of draws."""
    rng = random.Random(seed)
    counters = [0] * QUEUES
    state = {w: {"next": 0, "end": 0, "started": False} for w in range(waves)}
    live = list(range(waves))
    got = []
    while live:
        w = rng.choice(live)
        s = state[w]
        if s["next"] >= s["end"]:
            if first_static and not s["started"]:
                s["started"] = True
                run = w
            else:
                j = queue_of(w, waves)
                run = queue_run(w, waves, int(first_static), counters[j])
                counters[j] += 1
            s["next"], s["end"] = run * grab, run * grab + grab
        t = s["next"]
        s["next"] += 1
        if t >= ntiles:
            live.remove(w)
        else:
            got.append(t)
    return got


CASES = [(w, g, n, fs) for w in (1, 3, 15, 16, 17, 64, 100) for g in (1, 2, 5)
         for n in (0, 1, 7, 64, 333) for fs in (False, True)]


@pytest.mark.parametrize("waves,grab,ntiles,first_static", CASES)
def test_every_tile_exactly_once(hc, waves, grab, ntiles, first_static):
    for seed in range(3):
        got = play(hc.blh_queue_of, hc.blh_queue_run, waves, grab, ntiles, first_static, seed)
        assert sorted(got) == list(range(ntiles))


def test_queues_match_the_rule(hc):
    for waves in (1, 5, 16, 40):
        q = min(QUEUES, waves)
        for w in range(waves):
            assert hc.blh_queue_of(w, waves) == w % q
            for taken in (0, 1, 7):
                assert hc.blh_queue_run(w, waves, 1, taken) == waves + w % q + q * taken
                assert hc.blh_queue_run(w, waves, 0, taken) == w % q + q * taken


def test_a_queue_without_waves_is_caught():
    """All 16 queues with fewer waves: queues 3..15 are never drawn from."""
    def q_of(w, waves):
        return w % QUEUES

    def q_run(w, waves, fs, taken):
        return (waves if fs else 0) + w % QUEUES + QUEUES * taken

    got = play(q_of, q_run, 3, 1, 100, True, 0)
    assert sorted(got) != list(range(100))
