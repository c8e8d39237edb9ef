"""bw_bfs_run_static, the allocation-free form of bw_bfs_run, against
bw_bfs_run itself through the C ABI: identical depths, per-level counts
and level counts on two calls in a row (the second on a smaller graph, so
stale counters of the first would show), for every kernel family and
reduction, plus a high-diameter grid that uses many counters."""
import ctypes as C

import numpy as np
import pytest

from belewitte_amd import _native as N
from belewitte_amd import bfs, device, graph

pytestmark = pytest.mark.gpu


def _run(fn, dg, root, kernel, red, max_levels):
    lib = N.hip_lib()
    V = dg.vertex_count
    levels = device.DeviceArray(max(V, 1), np.int32)
    counts = np.zeros(max_levels, np.uint32)
    n = C.c_int64(0)
    sched = (C.c_int32 * 1)(kernel)
    cfg = N.launch()
    if fn == "static":
        st = lib.bw_bfs_run_static(dg.ref, sched, 1, red, root, levels.ptr, counts.ctypes.data_as(N.u32p),
                                   max_levels, C.byref(n), C.byref(cfg))
    else:
        st = lib.bw_bfs_run(dg.ref, sched, 1, red, root, levels.ptr, counts.ctypes.data_as(N.u32p), None,
                            max_levels, C.byref(n), C.byref(cfg))
    N.check_hip(st)
    out = levels.download()[:V]
    levels.free()
    return out, counts[:n.value].copy()


@pytest.mark.parametrize("kernel", range(8))
def test_static_matches_allocating_run_twice(gpu, kernel):
    graphs = [graph.rmat(13, 8, seed=21), graph.rmat(9, 4, seed=22)]  # second call on a smaller graph
    for red in range(4):
        for g in graphs:
            dg = device.DeviceGraph(g)
            dg.ensure(bfs.REPRESENTATION[kernel])
            root = g.max_out_degree_vertex()
            want, wc = _run("alloc", dg, root, kernel, red, 4096)
            got, gc = _run("static", dg, root, kernel, red, 4096)
            assert np.array_equal(got, want), (kernel, red, g.name)
            assert np.array_equal(gc, wc), (kernel, red, g.name)
            dg.free()


def test_static_many_levels_and_limits(gpu):
    lib = N.hip_lib()
    g = graph.grid2d(64, 48)
    dg = device.DeviceGraph(g)
    want, wc = _run("alloc", dg, 0, 4, 2, 1024)
    got, gc = _run("static", dg, 0, 4, 2, 1024)
    assert gc.size == wc.size == 64 + 48 - 1 and np.array_equal(got, want) and np.array_equal(gc, wc)
    # too few levels: the same limit error as bw_bfs_run
    levels = device.DeviceArray(g.vertex_count, np.int32)
    n = C.c_int64(0)
    sched = (C.c_int32 * 1)(4)
    st = lib.bw_bfs_run_static(dg.ref, sched, 1, 0, 0, levels.ptr, None, 10, C.byref(n), None)
    assert st == 5 and n.value == 10
    # beyond the static scratch: refused before any launch
    st = lib.bw_bfs_run_static(dg.ref, sched, 1, 0, 0, levels.ptr, None, 65537, C.byref(n), None)
    assert st == 1 and b"static scratch" in lib.bw_hip_last_error()
    # the run after the refusals is still exact
    got, gc = _run("static", dg, 0, 4, 2, 1024)
    assert np.array_equal(got, want) and np.array_equal(gc, wc)
    levels.free()
    dg.free()
