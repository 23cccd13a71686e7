"""The kernels' per-lane code under AddressSanitizer and UBSan, on the CPU.

tests/asan/lane_bounds.cpp builds bl_device.h for the host (through
bl_hostcheck.cpp) and runs load4, ByteStream, BlockStream, stripe_word, the
bit-plane build, the prefix/suffix trim, every lane kind (registers, 32-row
units, band, Peq search, wavefront) and packed sets over strings of 0-300
bytes placed at every alignment in heap blocks that end exactly at the end of
the last aligned unit holding a string byte. A read past (or before) those
units, or undefined behaviour, aborts it. The negative control reads the
dword after a string's last one and must be caught. No GPU needed.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import PKG, ROOT

SRC = os.path.join(ROOT, "tests", "asan", "lane_bounds.cpp")
CSRC = os.path.join(PKG, "csrc")


@pytest.fixture(scope="module")
def lane_bounds(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path_factory.mktemp("asan") / "lane_bounds")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-static-libasan",
           "-I", CSRC, SRC, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def _run(exe, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=600, env=env)


def test_lane_code_stays_in_its_aligned_units(lane_bounds):
    r = _run(lane_bounds)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "lane bounds: ok" in r.stdout


def test_a_read_past_the_last_dword_is_caught(lane_bounds):
    r = _run(lane_bounds, "overread")
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr
