"""Graph substrate (libbwgraph): generators, CSR invariants, `.graph` I/O."""
import os
import struct
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from belewitte_amd import graph
from graphfiles import duplicate_edge_graph

ROOT = Path(__file__).resolve().parents[1]


def _csr_ok(off, col, V):
    assert off[0] == 0 and off[-1] == col.size
    assert np.all(np.diff(off.astype(np.int64)) >= 0)
    assert col.size == 0 or col.max() < V
    for v in range(V):
        row = col[off[v]:off[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0), "rows sorted and unique"


def test_rmat_structure_and_determinism():
    g = graph.rmat(10, 8, seed=4)
    assert g.vertex_count == 1024 and 0 < g.edge_count <= 8 * 1024
    off, col = g.offsets(), g.cols()
    _csr_ok(off, col, g.vertex_count)
    src = np.repeat(np.arange(g.vertex_count), np.diff(off).astype(np.int64))
    assert not np.any(src == col), "self-loops removed"
    # reverse CSR is the transpose
    roff, rcol = g.offsets(True), g.cols(True)
    _csr_ok(roff, rcol, g.vertex_count)
    fwd = set(zip(src.tolist(), col.tolist()))
    rsrc = np.repeat(np.arange(g.vertex_count), np.diff(roff).astype(np.int64))
    assert fwd == set(zip(rcol.tolist(), rsrc.tolist()))
    h = graph.rmat(10, 8, seed=4)
    assert np.array_equal(h.cols(), col) and np.array_equal(h.offsets(), off)
    assert not np.array_equal(graph.rmat(10, 8, seed=5).cols()[:100], col[:100])


def test_rmat_quadrant_skew():
    # With a+b = 0.76 of the draws the source's top bit is 0.
    g = graph.rmat(14, 16, seed=1)
    src = np.repeat(np.arange(g.vertex_count), np.diff(g.offsets()).astype(np.int64))
    top = (src >> 13) & 1
    frac0 = 1.0 - top.mean()
    assert 0.72 < frac0 < 0.82
    dst_top = (g.cols() >> 13) & 1
    assert 0.72 < 1.0 - dst_top.mean() < 0.82


def test_generators_independent_of_thread_count():
    code = ("import sys,hashlib; sys.path.insert(0, %r); from belewitte_amd import graph; "
            "g = graph.rmat(12, 8, seed=9); e = graph.erdos_renyi(5000, 20000, seed=2); "
            "print(hashlib.sha1(g.cols().tobytes() + g.offsets(True).tobytes() + e.cols().tobytes()).hexdigest())"
            % str(ROOT))
    outs = set()
    for t in ("1", "3"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                                check=True).stdout.strip())
    assert len(outs) == 1


def test_undirected_is_symmetric_and_shared():
    g = graph.rmat(9, 8, seed=2, undirected=True)
    assert g.undirected
    off, col = g.offsets(), g.cols()
    src = np.repeat(np.arange(g.vertex_count), np.diff(off).astype(np.int64))
    pairs = set(zip(src.tolist(), col.tolist()))
    assert all((b, a) in pairs for a, b in pairs)
    assert np.array_equal(g.offsets(True), off)


def test_erdos_renyi_degree():
    n, m = 20000, 80000
    g = graph.erdos_renyi(n, m, seed=3, undirected=True)
    # ~2m directed edges (few duplicates / self-loops at this density)
    assert 0.99 * 2 * m < g.edge_count <= 2 * m
    _csr_ok(g.offsets(), g.cols(), n)


def test_grid_and_rewiring():
    W, H = 30, 20
    g = graph.grid2d(W, H, rewire=0.0)
    M = (W - 1) * H + (H - 1) * W
    assert g.edge_count == 2 * M
    deg = g.out_degrees().reshape(H, W)
    assert deg[1:-1, 1:-1].min() == 4 and deg[0, 0] == 2
    r = graph.grid2d(W, H, rewire=0.05, seed=3)
    off, col = r.offsets(), r.cols()
    src = np.repeat(np.arange(W * H), np.diff(off).astype(np.int64))
    lattice = np.abs(src.astype(np.int64) - col.astype(np.int64))
    nonlocal_edges = np.sum(~((lattice == 1) | (lattice == W))) // 2
    assert nonlocal_edges == pytest.approx(round(0.05 * M), abs=3)
    assert r.edge_count == 2 * M  # rewiring keeps the edge count (no duplicates made)


def test_closed_form_generators():
    c = graph.chain(5)
    assert c.vertex_count == 6 and c.cols().tolist() == [1, 2, 3, 4, 5]
    s = graph.star(4)
    assert s.out_degrees().tolist() == [4, 0, 0, 0, 0]
    m = graph.mesh(3, 4)
    assert m.vertex_count == 12 and m.edge_count == 2 * 3 * 4 - 3 - 4


def test_from_edges_dedup_and_self_loops():
    g = graph.from_edges(4, [0, 0, 1, 2, 2], [1, 1, 1, 3, 0])
    assert g.edge_count == 3  # duplicate 0->1 merged, self-loop 1->1 dropped
    k = graph.from_edges(4, [1], [1], keep_self_loops=True)
    assert k.edge_count == 1
    with pytest.raises(Exception):
        graph.from_edges(3, [0], [5])


def test_max_out_degree_tie_lowest_id():
    g = graph.from_edges(5, [3, 3, 1, 1, 0], [0, 1, 2, 3, 1])
    assert g.max_out_degree_vertex() == 1


def test_degree_stats_match_numpy():
    g = graph.rmat(9, 8, seed=6)
    out = g.out_degrees().astype(np.float64)
    st = g.degree_stats("out")
    assert st["min"] == out.min() and st["max"] == out.max()
    assert st["mean"] == pytest.approx(out.mean())
    assert st["median"] == pytest.approx(np.median(out))
    assert st["stddev"] == pytest.approx(out.std(ddof=1))
    ab = g.degree_stats("abs")
    assert ab["mean"] == pytest.approx(2 * out.mean())


def test_graph_file_roundtrip(tmp_path):
    for g in (graph.rmat(8, 4, seed=1), graph.rmat(8, 4, seed=1, undirected=True)):
        p = tmp_path / "g.graph"
        g.save(p)
        raw = p.read_bytes()
        hdr = struct.unpack("<6I2Q", raw[:40])
        assert hdr[:6] == (1, 0, 0, int(g.undirected), 4, 4)
        assert hdr[6] == g.vertex_count and hdr[7] == g.edge_count
        h = graph.load(p)
        assert h.undirected == g.undirected
        assert np.array_equal(h.offsets(), g.offsets()) and np.array_equal(h.cols(), g.cols())
        assert np.array_equal(h.offsets(True), g.offsets(True)) and np.array_equal(h.cols(True), g.cols(True))


def test_graph_file_v0_and_errors(tmp_path):
    # v0: int32 [undirected, V, E], offsets, edges (+ reverse when directed)
    V, E = 3, 2
    body = [0, V, E, 0, 1, 2, 2, 1, 2, 0, 0, 1, 2, 0, 1]
    p = tmp_path / "v0.graph"
    p.write_bytes(struct.pack("<%di" % len(body), *body))
    g = graph.load(p)
    assert (g.vertex_count, g.edge_count, g.undirected) == (3, 2, False)
    assert g.cols().tolist() == [1, 2] and g.cols(True).tolist() == [0, 1]
    p.write_bytes(struct.pack("<%di" % (len(body) - 1), *body[:-1]))
    with pytest.raises(Exception, match="Invalid file size"):
        graph.load(p)
    with pytest.raises(Exception):
        graph.load(tmp_path / "missing.graph")


def test_reorder_by_degree():
    g = graph.rmat(9, 8, seed=12)
    r, perm = g.reorder_by_degree()
    d = r.out_degrees()
    assert np.all(np.diff(d.astype(np.int64)) <= 0)
    assert r.edge_count == g.edge_count
    assert np.array_equal(np.sort(perm), np.arange(g.vertex_count))


def test_reorder_keeps_duplicate_edges(tmp_path):
    """The degree relabelling is a pure permutation: a file with duplicate
    edges keeps its edge count and its edge multiset through load and
    reorder_by_degree, in both stored directions."""
    p = tmp_path / "dup.graph"
    src, dst = duplicate_edge_graph(p)
    g = graph.load(p)
    assert g.edge_count == src.size
    r, perm = g.reorder_by_degree()
    assert r.edge_count == g.edge_count
    want = sorted(zip(perm[src].tolist(), perm[dst].tolist()))
    s, c = r.edges(False)
    assert sorted(zip(s.tolist(), c.tolist())) == want
    rd, rs = r.edges(True)
    assert sorted(zip(rs.tolist(), rd.tolist())) == want
    assert np.all(np.diff(r.out_degrees().astype(np.int64)) <= 0)
