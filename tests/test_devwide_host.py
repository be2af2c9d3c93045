"""Lock discipline of device-wide runtime calls against stream capture (host-side check).

Contexts may be driven from several host threads, each capturing its own stream into a HIP
graph (gf_use_graphs).  hipMalloc, hipFree, synchronous copies, null-stream memsets and
synchronisations, stream creation and destruction and graph-exec destruction must not run on
one thread while another thread is capturing.  The library routes every such call through the
wrappers between the DEVWIDE markers of gf_context.cpp, which hold one process-wide
shared_mutex shared, while a capture holds it exclusively.  This test checks, on the sources,
that no raw call escapes the wrappers and that the capture path takes the lock exclusively in
thread-local capture mode.
"""
import glob
import os
import re

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "geophysicalflows.jl_amd", "csrc")
RAW = re.compile(r"\b(hipMalloc|hipFree|hipHostMalloc|hipHostFree|hipMemcpy|hipMemset|hipStreamCreate\w*|hipStreamDestroy|"
                 r"hipGraphExecDestroy|hipStreamBeginCapture|hipStreamEndCapture|hipGraphInstantiate|hipDeviceSynchronize)\s*\(")
NULL_SYNC = re.compile(r"hipStreamSynchronize\s*\(\s*(nullptr|0|NULL)\s*\)")


def _code(line):
    return line.split("//", 1)[0]


def test_raw_device_wide_calls_only_inside_the_wrappers():
    found_markers = False
    for path in sorted(glob.glob(os.path.join(CSRC, "*"))):
        if not path.endswith((".cpp", ".hip", ".hpp")):
            continue
        inside = False
        for no, line in enumerate(open(path), 1):
            if "DEVWIDE-BEGIN" in line:
                inside, found_markers = True, True
                continue
            if "DEVWIDE-END" in line:
                inside = False
                continue
            code = _code(line)
            if not inside:
                assert not RAW.search(code), f"{os.path.basename(path)}:{no}: raw device-wide call outside dw::  {line.strip()}"
                assert not NULL_SYNC.search(code), f"{os.path.basename(path)}:{no}: null-stream sync outside dw::"
    assert found_markers


def test_wrappers_lock_shared_and_capture_locks_exclusive():
    src = open(os.path.join(CSRC, "gf_context.cpp")).read()
    block = src.split("DEVWIDE-BEGIN", 1)[1].split("DEVWIDE-END", 1)[0]
    funcs, cur = [], None
    for line in block.splitlines():
        m = re.match(r"^(?:hipError_t|hipGraphExec_t) (\w+)\(", line)
        if m:
            cur = [m.group(1), line]
            funcs.append(cur)
        elif cur is not None:
            cur[1] += "\n" + line
            if line == "}":
                cur = None
    names = {n for n, _ in funcs}
    assert {"malloc", "free", "copy", "zero_sync", "stream_create", "stream_destroy", "exec_destroy", "capture"} <= names
    for name, body in funcs:
        if name == "capture":
            assert "std::unique_lock<std::shared_mutex> l(g_devwide)" in body
            assert "hipStreamCaptureModeThreadLocal" in body and "hipStreamCaptureModeRelaxed" not in body
            # the lock is taken before the capture begins and held through instantiation
            assert body.index("unique_lock") < body.index("hipStreamBeginCapture") < body.index("hipGraphInstantiate")
        else:
            assert "SL l(g_devwide)" in body, name
    # every graph of the library is captured through dw::capture
    assert src.count("dw::capture(") >= 2
