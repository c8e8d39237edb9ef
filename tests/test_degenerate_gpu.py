"""Degenerate graphs through the native BFS engine, the multi-source BFS
and the fused PageRank: a single vertex, vertices without any edge, a
single edge, and two components.  Depths against the oracle's queue BFS
(bit for bit); ranks against the FP64 oracle recurrence."""
import numpy as np
import pytest

from belewitte_amd import bfs_native, device, graph, pagerank, switch
from oracle import binding as O

pytestmark = pytest.mark.gpu


def degenerate_graphs():
    empty = np.zeros(0, np.uint32)
    g = graph.from_edges(1, empty, empty)
    g.name = "one-vertex"
    yield g
    g = graph.from_edges(5, empty, empty)
    g.name = "five-isolated"
    yield g
    g = graph.from_edges(3, np.array([0], np.uint32), np.array([1], np.uint32))
    g.name = "one-edge"
    yield g
    # a directed chain 0 -> ... -> 9 and a star 10 -> 11..19
    s = np.array(list(range(9)) + [10] * 9, np.uint32)
    d = np.array(list(range(1, 10)) + list(range(11, 20)), np.uint32)
    g = graph.from_edges(20, s, d)
    g.name = "chain-and-star"
    yield g


@pytest.mark.parametrize("policy", ["td", "bu", "bu-wave", "learned"])
def test_engine_degenerate(gpu, policy):
    model = switch.Model.load(switch.MODEL_PATH).w
    for g in degenerate_graphs():
        csr = O.CSR.from_graph(g)
        dg = device.DeviceGraph(g)
        eng = bfs_native.BFSEngine(dg)
        eng.set_model(model)
        for root in range(g.vertex_count):
            r = eng.run(root, "learned", "td") if policy == "learned" else eng.run(root, "fixed", policy)
            want, _ = O.bfs_queue(csr, root)
            assert np.array_equal(r.levels, want), (g.name, root, policy)
        eng.close()
        dg.free()


@pytest.mark.parametrize("td", [0, 1])
def test_multi_source_degenerate(gpu, td):
    for g in degenerate_graphs():
        csr = O.CSR.from_graph(g)
        dg = device.DeviceGraph(g)
        ms = bfs_native.MultiSourceBFS(dg)
        ms.set_tuning(32, td, 16)
        src = np.arange(min(g.vertex_count, 64), dtype=np.uint32)
        depths, _, _ = ms.run(src)
        for i, s in enumerate(src):
            want, _ = O.bfs_queue(csr, int(s))
            assert np.array_equal(depths[i], want), (g.name, int(s), td)
        ms.close()
        dg.free()


@pytest.mark.parametrize("mode", ["drop", "uniform"])
@pytest.mark.parametrize("width", [0, 2])
def test_fused_pagerank_degenerate(gpu, mode, width):
    for g in degenerate_graphs():
        csr = O.CSR.from_graph(g)
        fp = pagerank.FusedPageRank.for_graph(g, max_iterations=10, reorder=True)
        fp.tune(0, 0, 1, width, 16)
        ranks, resid, _ = fp.run(10, mode)
        want = O.pagerank_f64(csr, 10, mode)
        assert float(np.max(np.abs(ranks - want))) <= 1e-6, (g.name, mode, width)
        assert np.all(np.isfinite(resid))
        fp.close()


@pytest.mark.parametrize("offset_bytes", [4, 8])
def test_edge_lists_match_the_csr(gpu, offset_bytes):
    """The edge lists built on the device (bw_build_edge_lists) hold every
    CSR edge in CSR order: graphs with empty rows, one very long row and a
    chain, both offset widths, both directions."""
    graphs = [graph.star(70000), graph.chain(5000), graph.rmat(14, 8, seed=41),
              graph.from_edges(9000, np.array([8999, 8999, 3], np.uint32), np.array([0, 1, 2], np.uint32))]
    for g in graphs:
        dg = device.DeviceGraph(g, offset_bytes=offset_bytes, reps=device.REPRESENTATIONS)
        try:
            for rev in (False, True):
                s, d = g.edges(reverse=rev)
                ki, ko, kp = ("rel_in", "rel_out", "rsel") if rev else ("el_in", "el_out", "sel")
                a_in = dg.array(ki).download()[:g.edge_count]
                a_out = dg.array(ko).download()[:g.edge_count]
                pairs = dg.array(kp).download()[:2 * g.edge_count]
                assert np.array_equal(a_in, s) and np.array_equal(a_out, d), (g.name, rev)
                assert np.array_equal(pairs[0::2], s) and np.array_equal(pairs[1::2], d), (g.name, rev)
        finally:
            dg.free()
