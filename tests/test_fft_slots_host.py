"""Host checks of index arithmetic shared by the kernels (no GPU needed): the FFT core's split
LDS slots and the ring-forcing box rows.

The kernels address the padded LDS buffers as a per-thread base slot plus compile-time offsets
(``padg<G>(a + b) = padg<G>(a) + padg<G>(b)`` when ``a mod G + b mod G < G``, for the padding
period G of each exchange).  tests/host/fft_slots_check.cpp compares that form with the direct
slot for every point of every plan, both point sizes and every period, exhaustively.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "geophysicalflows.jl_amd", "csrc")
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.mark.skipif(HIPCC is None, reason="hipcc is not installed")
def test_split_slots_match_direct_slots(tmp_path):
    exe = tmp_path / "fft_slots_check"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "fft_slots_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("total 0")


@pytest.mark.skipif(HIPCC is None, reason="hipcc is not installed")
def test_ring_box_rows_are_one_to_one(tmp_path):
    exe = tmp_path / "ring_box_check"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I", CSRC, os.path.join(HERE, "host", "ring_box_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("mismatches 0")
