"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker -- never by the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

UNVISITED = 2147483647
BFS_VARIANTS = ("edge-list", "rev-edge-list", "struct-edge-list", "rev-struct-edge-list",
                "vertex-push", "vertex-pull", "vertex-push-warp", "vertex-pull-warp")
PR_VARIANTS = ("edge-list", "rev-edge-list", "struct-edge-list", "rev-struct-edge-list",
               "vertex-push", "vertex-pull", "vertex-push-warp", "vertex-pull-warp",
               "vertex-pull-nodiv", "vertex-pull-warp-nodiv")
MODES = {"drop": 0, "uniform": 1}


class OracleGraph(C.Structure):
    _fields_ = [("V", C.c_uint64), ("E", C.c_uint64),
                ("off", C.POINTER(C.c_uint64)), ("col", C.POINTER(C.c_uint32)),
                ("roff", C.POINTER(C.c_uint64)), ("rcol", C.POINTER(C.c_uint32))]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        G = C.POINTER(OracleGraph)
        L.oracle_bfs_queue.restype = C.c_int64
        L.oracle_bfs_queue.argtypes = [G, C.c_uint32, C.c_void_p]
        L.oracle_bfs_level.restype = C.c_uint32
        L.oracle_bfs_level.argtypes = [G, C.c_int, C.c_void_p, C.c_int32, C.c_uint32, C.c_uint32]
        L.oracle_bfs_run.restype = C.c_int64
        L.oracle_bfs_run.argtypes = [G, C.c_uint32, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32,
                                     C.c_void_p, C.c_void_p, C.c_int64]
        L.oracle_bfs_vertex_push_omp.restype = C.c_int64
        L.oracle_bfs_vertex_push_omp.argtypes = [G, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int64, C.c_int,
                                                 C.POINTER(C.c_double)]
        L.oracle_pr_out_degrees.restype = None
        L.oracle_pr_out_degrees.argtypes = [G, C.c_void_p]
        L.oracle_pagerank_f32.restype = None
        L.oracle_pagerank_f32.argtypes = [G, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_void_p, C.c_void_p]
        L.oracle_pagerank_f64.restype = None
        L.oracle_pagerank_f64.argtypes = [G, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.oracle_pagerank_binned_fixed.restype = None
        L.oracle_pagerank_binned_fixed.argtypes = [G, C.c_int, C.c_int, C.c_void_p]
        L.oracle_pagerank_pull_omp.restype = None
        L.oracle_pagerank_pull_omp.argtypes = [G, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        L.oracle_pagerank_bucket_order.restype = None
        L.oracle_pagerank_bucket_order.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        _lib = L
    return _lib


class CSR:
    """Holds the arrays an OracleGraph points at (keeps them alive)."""

    def __init__(self, V, off, col, roff=None, rcol=None):
        self.V = int(V)
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.roff = self.off if roff is None else np.ascontiguousarray(roff, dtype=np.uint64)
        self.rcol = self.col if rcol is None else np.ascontiguousarray(rcol, dtype=np.uint32)
        self.E = int(self.col.size)
        self.s = OracleGraph(self.V, self.E,
                             self.off.ctypes.data_as(C.POINTER(C.c_uint64)),
                             self.col.ctypes.data_as(C.POINTER(C.c_uint32)),
                             self.roff.ctypes.data_as(C.POINTER(C.c_uint64)),
                             self.rcol.ctypes.data_as(C.POINTER(C.c_uint32)))

    @classmethod
    def from_graph(cls, g):
        if g.undirected:
            c = cls(g.vertex_count, g.offsets(), g.cols())
        else:
            c = cls(g.vertex_count, g.offsets(), g.cols(), g.offsets(True), g.cols(True))
        c.graph = g  # the arrays are views into the graph's memory
        return c

    @property
    def ref(self):
        return C.byref(self.s)


def bfs_queue(csr: CSR, root: int):
    depth = np.empty(max(csr.V, 1), dtype=np.int32)
    n = lib().oracle_bfs_queue(csr.ref, root, depth.ctypes.data)
    return depth[:csr.V], int(n)


def bfs_level(csr: CSR, variant: int, levels: np.ndarray, depth: int, warp=32, chunk=32) -> int:
    assert levels.dtype == np.int32 and levels.flags.c_contiguous
    return int(lib().oracle_bfs_level(csr.ref, variant, levels.ctypes.data, depth, warp, chunk))


def bfs_run(csr: CSR, root: int, schedule, warp=32, chunk=32, max_levels=None):
    sched = np.ascontiguousarray(schedule, dtype=np.int32)
    max_levels = max_levels or csr.V + 2
    levels = np.empty(max(csr.V, 1), dtype=np.int32)
    counts = np.zeros(max_levels, dtype=np.uint32)
    n = lib().oracle_bfs_run(csr.ref, root, sched.ctypes.data, sched.size, warp, chunk,
                             levels.ctypes.data, counts.ctypes.data, max_levels)
    if n < 0:
        raise RuntimeError("oracle BFS exceeded max_levels")
    return levels[:csr.V], counts[:n].copy()


def bfs_vertex_push_omp(csr: CSR, root: int, threads: int = 0, max_levels=None, timed: bool = False):
    """Levels and per-level counts; with timed=True also the seconds of the
    initialisation and level loop (the reference's computation scope)."""
    max_levels = max_levels or csr.V + 2
    levels = np.empty(max(csr.V, 1), dtype=np.int32)
    counts = np.zeros(max_levels, dtype=np.uint32)
    sec = C.c_double(0.0)
    n = lib().oracle_bfs_vertex_push_omp(csr.ref, root, levels.ctypes.data, counts.ctypes.data,
                                         max_levels, threads, C.byref(sec))
    if timed:
        return levels[:csr.V], counts[:n].copy(), sec.value
    return levels[:csr.V], counts[:n].copy()


def out_degrees(csr: CSR) -> np.ndarray:
    deg = np.empty(max(csr.V, 1), dtype=np.uint32)
    lib().oracle_pr_out_degrees(csr.ref, deg.ctypes.data)
    return deg[:csr.V]


def pagerank_f32(csr: CSR, variant: int, iterations: int, mode: str = "drop", warp: int = 32):
    ranks = np.empty(max(csr.V, 1), dtype=np.float32)
    resid = np.zeros(max(iterations, 1), dtype=np.float64)
    lib().oracle_pagerank_f32(csr.ref, variant, iterations, MODES[mode], warp, ranks.ctypes.data,
                              resid.ctypes.data)
    return ranks[:csr.V], resid[:iterations]


def pagerank_f64(csr: CSR, iterations: int, mode: str = "drop", nodiv_quirk: bool = False):
    x = np.empty(max(csr.V, 1), dtype=np.float64)
    lib().oracle_pagerank_f64(csr.ref, iterations, MODES[mode], int(nodiv_quirk), x.ctypes.data)
    return x[:csr.V]


def pagerank_binned_fixed(csr: CSR, iterations: int, nodiv_quirk: bool = False) -> np.ndarray:
    """DROP-mode ranks with the binned iteration's fixed-point row sums
    (oracle_pagerank_binned_fixed)."""
    r = np.empty(max(csr.V, 1), dtype=np.float32)
    lib().oracle_pagerank_binned_fixed(csr.ref, iterations, int(nodiv_quirk), r.ctypes.data)
    return r[:csr.V]


def pagerank_pull_omp(csr: CSR, iterations: int, threads: int = 0, timed: bool = False):
    """Ranks; with timed=True also the seconds of the degrees and iterations
    (the reference's computation scope)."""
    ranks = np.empty(max(csr.V, 1), dtype=np.float32)
    sec = C.c_double(0.0)
    lib().oracle_pagerank_pull_omp(csr.ref, iterations, ranks.ctypes.data, threads, C.byref(sec))
    if timed:
        return ranks[:csr.V], sec.value
    return ranks[:csr.V]


def bucket_order(ranks: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(ranks, dtype=np.float32)
    out = np.empty(max(r.size, 1), dtype=np.uint64)
    lib().oracle_pagerank_bucket_order(r.ctypes.data, r.size, out.ctypes.data)
    return out[:r.size]
