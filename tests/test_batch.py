"""Host side of the batched BFS: the disjoint union of many graphs and the
seeded batch of small R-MAT graphs (no GPU needed)."""
import numpy as np

from belewitte_amd import batch, graph
from oracle import binding as O


def test_union_is_disjoint_and_preserves_each_graph():
    gs = [graph.rmat_n(100, 400, seed=1), graph.chain(7), graph.star(5, undirected=True), graph.rmat_n(60, 90, seed=2)]
    u, base = batch.union(gs)
    assert base == [0, 100, 108, 114, 174]
    assert u.vertex_count == 174
    assert u.edge_count == sum(g.edge_count for g in gs)
    us, ud = u.edges()
    for i, g in enumerate(gs):
        s, d = g.edges()
        m = (us >= base[i]) & (us < base[i + 1])
        assert np.all((ud[m] >= base[i]) & (ud[m] < base[i + 1]))  # no edge leaves its component
        got = sorted(zip((us[m] - base[i]).tolist(), (ud[m] - base[i]).tolist()))
        assert got == sorted(zip(s.tolist(), d.tolist()))


def test_multi_root_depths_are_per_component_depths():
    """The property the many-graph batch relies on, checked on the oracle:
    BFS from all roots of a disjoint union equals each graph's own BFS."""
    gs = [graph.rmat_n(200, 1200, seed=5), graph.grid2d(9, 7), graph.chain(30)]
    u, base = batch.union(gs)
    # a virtual super-root joined to every root reproduces the multi-root search
    roots = [base[i] + g.max_out_degree_vertex() for i, g in enumerate(gs)]
    s, d = u.edges()
    V = u.vertex_count
    sup = graph.from_edges(V + 1, np.concatenate([s, np.full(len(roots), V, np.uint32)]),
                           np.concatenate([d, np.asarray(roots, np.uint32)]))
    dep, _ = O.bfs_queue(O.CSR.from_graph(sup), V)
    multi = np.where(dep[:V] == O.UNVISITED, O.UNVISITED, dep[:V] - 1)
    for i, g in enumerate(gs):
        want, _ = O.bfs_queue(O.CSR.from_graph(g), roots[i] - base[i])
        assert np.array_equal(multi[base[i]:base[i + 1]], want)


def test_small_rmat_batch_is_seeded_and_in_range():
    a = batch.small_rmat_batch(12, seed=3, log2_min=6, log2_max=9, ef_min=2, ef_max=5)
    b = batch.small_rmat_batch(12, seed=3, log2_min=6, log2_max=9, ef_min=2, ef_max=5)
    assert [g.vertex_count for g in a] == [g.vertex_count for g in b]
    assert [g.edge_count for g in a] == [g.edge_count for g in b]
    for g in a:
        assert 2 ** 6 - 1 <= g.vertex_count <= 2 ** 9 + 1
        assert g.edge_count <= 5 * g.vertex_count
    c = batch.small_rmat_batch(12, seed=4, log2_min=6, log2_max=9, ef_min=2, ef_max=5)
    assert [g.vertex_count for g in a] != [g.vertex_count for g in c]
