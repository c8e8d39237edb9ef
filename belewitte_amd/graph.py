"""Host graphs: generators, CSR views and the `.graph` file format.

A :class:`Graph` owns a native ``bw_graph`` (libbwgraph.so) and exposes its
CSR arrays as zero-copy numpy views.  The layout follows the reference's
``Graph<V,E>`` (utils/Graph.hpp:784-1063): a forward CSR of out-edges and,
for directed graphs, a reverse CSR of in-edges; an undirected graph stores
one symmetric CSR serving both directions.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

# R-MAT quadrant probabilities of the Graph500 generator.
RMAT_ABC = (0.57, 0.19, 0.19)


class Graph:
    """Owning wrapper of a native graph."""

    def __init__(self, handle: int, name: str = "graph"):
        if not handle:
            raise ValueError("null graph handle")
        self._h = C.c_void_p(handle)
        self.name = name
        lib = N.graph_lib()
        self.vertex_count = int(lib.bw_graph_vertex_count(self._h))
        self.edge_count = int(lib.bw_graph_edge_count(self._h))
        self.undirected = bool(lib.bw_graph_is_undirected(self._h))

    # ------------------------------------------------------------ lifetime
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            N.graph_lib().bw_graph_free(self._h)
            self._h = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def __repr__(self) -> str:
        kind = "undirected" if self.undirected else "directed"
        return f"Graph({self.name!r}, V={self.vertex_count}, E={self.edge_count}, {kind})"

    # --------------------------------------------------------------- views
    def offsets(self, reverse: bool = False) -> np.ndarray:
        """uint64 CSR offsets [V+1] (view, valid while the graph lives)."""
        p = N.graph_lib().bw_graph_offsets(self._h, int(reverse))
        return np.ctypeslib.as_array(p, shape=(self.vertex_count + 1,))

    def cols(self, reverse: bool = False) -> np.ndarray:
        """uint32 CSR columns [E] (view)."""
        if self.edge_count == 0:
            return np.zeros(0, dtype=np.uint32)
        p = N.graph_lib().bw_graph_cols(self._h, int(reverse))
        return np.ctypeslib.as_array(p, shape=(self.edge_count,))

    def offsets32(self, reverse: bool = False) -> np.ndarray:
        out = np.empty(self.vertex_count + 1, dtype=np.uint32)
        N.check_graph(N.graph_lib().bw_graph_offsets32(
            self._h, int(reverse), out.ctypes.data_as(N.u32p)))
        return out

    def out_degrees(self) -> np.ndarray:
        return np.diff(self.offsets()).astype(np.uint32)

    def in_degrees(self) -> np.ndarray:
        return np.diff(self.offsets(True)).astype(np.uint32)

    def edges(self, reverse: bool = False):
        """EdgeList (in[], out[]) in the order of GraphLoader.hpp:134-143."""
        off = self.offsets(reverse)
        rows = np.repeat(np.arange(self.vertex_count, dtype=np.uint32), np.diff(off).astype(np.int64))
        return rows, self.cols(reverse).copy()

    def max_out_degree_vertex(self) -> int:
        return int(N.graph_lib().bw_graph_max_out_degree_vertex(self._h))

    def degree_stats(self, kind: str) -> dict:
        """utils/Graph.hpp degreeStatistics for kind in {'in','out','abs'}."""
        k = {"in": 0, "out": 1, "abs": 2}[kind]
        out = (C.c_double * 7)()
        N.check_graph(N.graph_lib().bw_graph_degree_stats(self._h, k, out))
        keys = ("min", "lower quantile", "mean", "median", "upper quantile", "max", "stddev")
        return dict(zip(keys, list(out)))

    def save(self, path: str) -> None:
        N.check_graph(N.graph_lib().bw_graph_save(self._h, str(path).encode()))

    def reorder_by_degree(self):
        h = C.c_void_p()
        perm = np.empty(self.vertex_count, dtype=np.uint32)
        N.check_graph(N.graph_lib().bw_graph_reorder_by_degree(
            C.byref(h), self._h, perm.ctypes.data_as(N.u32p)))
        return Graph(h.value, self.name + "-reordered"), perm


def _make(fn, *args, name: str) -> Graph:
    h = C.c_void_p()
    N.check_graph(fn(C.byref(h), *args))
    return Graph(h.value, name)


def rmat(scale: int, edge_factor: int = 16, seed: int = 1, undirected: bool = False,
         abc=RMAT_ABC) -> Graph:
    a, b, c = abc
    return _make(N.graph_lib().bw_graph_rmat, scale, edge_factor, a, b, c, seed, int(undirected),
                 name=f"rmat-s{scale}-ef{edge_factor}-{'u' if undirected else 'd'}-seed{seed}")


def rmat_n(vertex_count: int, edge_count: int, seed: int = 1, undirected: bool = False,
           abc=RMAT_ABC) -> Graph:
    a, b, c = abc
    return _make(N.graph_lib().bw_graph_rmat_n, vertex_count, edge_count, a, b, c, seed,
                 int(undirected), name=f"rmat-n{vertex_count}-m{edge_count}-seed{seed}")


def erdos_renyi(vertex_count: int, pairs: int, seed: int = 1, undirected: bool = True) -> Graph:
    return _make(N.graph_lib().bw_graph_erdos_renyi, vertex_count, pairs, seed, int(undirected),
                 name=f"er-n{vertex_count}-m{pairs}-seed{seed}")


def grid2d(width: int, height: int, rewire: float = 0.0, seed: int = 1) -> Graph:
    return _make(N.graph_lib().bw_graph_grid2d, width, height, rewire, seed,
                 name=f"grid-{width}x{height}-r{rewire}-seed{seed}")


def chain(n: int, undirected: bool = False) -> Graph:
    return _make(N.graph_lib().bw_graph_chain, n, int(undirected), name=f"chain-{n}")


def star(n: int, undirected: bool = False) -> Graph:
    return _make(N.graph_lib().bw_graph_star, n, int(undirected), name=f"star-{n}")


def mesh(height: int, width: int, undirected: bool = False) -> Graph:
    return _make(N.graph_lib().bw_graph_mesh, height, width, int(undirected),
                 name=f"mesh-{height}x{width}")


def from_edges(vertex_count: int, src, dst, undirected: bool = False,
               keep_self_loops: bool = False) -> Graph:
    s = np.ascontiguousarray(src, dtype=np.uint32)
    d = np.ascontiguousarray(dst, dtype=np.uint32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    return _make(N.graph_lib().bw_graph_from_edges, vertex_count, s.size,
                 s.ctypes.data_as(N.u32p), d.ctypes.data_as(N.u32p), int(undirected),
                 int(keep_self_loops), name="edges")


def load(path: str) -> Graph:
    return _make(N.graph_lib().bw_graph_load, str(path).encode(), name=str(path))


def hash64(seed: int, stream: int, counter: int) -> int:
    return int(N.graph_lib().bw_hash64(seed, stream, counter))


def random_sources(g: Graph, count: int, seed: int) -> list:
    """`count` distinct sources with out-degree >= 1, drawn by hash rank."""
    deg = g.out_degrees()
    cand = np.nonzero(deg > 0)[0]
    if cand.size == 0:
        return []
    keys = _vector_hash(seed, 7, cand)
    order = np.argsort(keys, kind="stable")
    return [int(cand[i]) for i in order[:count]]


def _vector_hash(seed: int, stream: int, ctr: np.ndarray) -> np.ndarray:
    # Same splitmix64 finaliser as bw_hash64, vectorised in numpy.
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (ctr.astype(np.uint64) + np.uint64(1))
             + np.uint64(0xD1B54A32D192ED03) * np.uint64(stream + 1))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


@dataclass
class GraphSpec:
    """Named synthetic input of the benchmark configurations."""
    kind: str
    params: dict

    def build(self) -> Graph:
        if self.kind == "rmat":
            return rmat(**self.params)
        if self.kind == "er":
            return erdos_renyi(**self.params)
        if self.kind == "grid":
            return grid2d(**self.params)
        raise ValueError(self.kind)
