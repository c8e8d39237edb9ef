"""The C ABI libraries load and export every symbol their headers declare.

CPU-only: no kernel is launched.  Argument validation that happens on the
host before any launch is exercised (it returns BW_HIP_ERR_ARG with a
message instead of touching the device).
"""
import ctypes as C
import re
from pathlib import Path

import pytest

from belewitte_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def _declared(header: str):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bw_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.mark.parametrize("header,loader,sigs", [
    ("belewitte_hip.h", N.hip_lib, N.hip_signatures),
    ("belewitte_graph.h", N.graph_lib, N.graph_signatures),
])
def test_every_declared_symbol_is_exported_and_bound(header, loader, sigs):
    names = _declared(header)
    assert len(names) > 10
    lib = loader()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in {header} but not exported"
    missing = [n for n in names if n not in sigs()]
    assert not missing, f"no ctypes signature for {missing}"


def test_device_count_without_gpu_is_ok():
    n = C.c_int32(-1)
    assert N.hip_lib().bw_hip_device_count(C.byref(n)) == 0
    assert n.value >= 0


def test_host_side_argument_checks():
    lib = N.hip_lib()
    cfg = N.launch()
    # NULL levels / frontier are rejected before any launch.
    st = lib.bw_bfs_vertex_push(10, None, 4, None, None, 0, None, 0, C.byref(cfg))
    assert st == 1 and b"device pointers" in lib.bw_hip_last_error()
    st = lib.bw_bfs_edge_list(0, None, None, None, 0, None, 0, C.byref(cfg))
    assert st == 1
    # Unknown variant / mode ids.
    g = N.BwGraphDev()
    assert lib.bw_pr_run(C.byref(g), 99, 1, 0, C.c_void_p(1), None, None) == 1
    assert b"unknown PageRank variant" in lib.bw_hip_last_error()
    assert lib.bw_pr_run(C.byref(g), 5, 1, 7, C.c_void_p(1), None, None) == 1
    assert lib.bw_bfs_level(C.byref(g), 42, 0, C.c_void_p(1), 0, C.c_void_p(1), None) == 1
    assert b"unknown BFS variant" in lib.bw_hip_last_error()
    assert lib.bw_bfs_run(None, None, 0, 0, 0, None, None, None, 0, None, None) == 1
    # The allocation-free loop refuses more levels than its static counters
    # hold, before touching the device.
    sched = (C.c_int32 * 1)(4)
    n = C.c_int64(0)
    assert lib.bw_bfs_run_static(C.byref(g), sched, 1, 0, 0, C.c_void_p(1), None, 65537, C.byref(n), None) == 1
    assert b"static scratch" in lib.bw_hip_last_error()


def test_version_string():
    v = N.hip_lib().bw_hip_version().decode()
    assert v.startswith("belewitte-amd")
