"""Host checks of bench.py's contract that need no GPU: the committed HBM-traffic summary names
the kernels of the shipped 2048^2 FP64 schedule (so the bench line's `traffic` field is filled),
the kernel trace summary times the same kernels, and a request for several GPUs is refused with
a message."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _shipped_2048_fp64():
    with open(os.path.join(ROOT, "geophysicalflows.jl_amd", "schedules.json")) as f:
        ents = json.load(f)["entries"]
    e = [x for x in ents if x["model"] == 0 and x["nx"] == 2048 and x["ny"] == 2048 and x["dtype"] == 0]
    assert len(e) == 1
    return tuple(e[0]["p"]), tuple(e[0]["split"])


def test_traffic_summary_covers_the_shipped_schedule():
    sched, splits = _shipped_2048_fp64()
    tr = bench.measured_traffic(sched, splits)
    assert set(tr) == set(bench.KERNELS)
    for k, b in tr.items():
        assert 50e6 < b < 200e6, (k, b)   # tens of MB per launch at 2048^2 FP64


def test_kernel_trace_summary_times_the_shipped_kernels():
    sched, splits = _shipped_2048_fp64()
    with open(os.path.join(ROOT, "profiles", "twodns_2048_fp64_etdrk4_kernel_trace.md")) as f:
        txt = f.read()
    for syms in bench.symbols(sched, splits).values():
        for s in syms:
            assert f"`{s}`" in txt, s


def test_several_gpus_are_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert "one GPU" in (r.stdout + r.stderr)
