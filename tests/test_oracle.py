"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

The reference ships no golden vectors, known-answer tests or fixtures for
this path (no tests at all; its correctness check is cross-implementation
output hashing, benchmark-analysis/ingest-src/Jobs.hs:230-241), so the oracle
is pinned by what the mathematics fixes:

BFS (every variant restatement, oracle_bfs_level / oracle_bfs_run):
  * an independent shortest-path routine (networkx when importable, else a
    dense Bellman-Ford in numpy) on random small graphs;
  * closed forms of the generators of evolve/gen-graph.cpp:72-123
    (chain depth i, star depth 1, mesh depth x + y from the corner);
  * the BFS invariants (d[root] = 0, d[v] <= d[u] + 1 along every edge,
    every reached vertex has an in-neighbour one level up, unreached
    vertices have no reached in-neighbour) -- a planted off-by-one in the
    depths fails them;
  * per-launch frontier counts = |{v : d[v] = k + 1}|, launches = depth + 1.
PageRank (oracle_pagerank_f64 / _f32):
  * dense-matrix power iteration in numpy (an independent formulation);
  * closed forms: a directed cycle keeps 1/N; a star's two-value solution;
    UNIFORM mode conserves total mass 1; DROP mode loses exactly the
    dangling mass;
  * the nodiv first-iteration quirk: undivided 1/N contributions;
  * FP32 restatement within FP32 rounding of the FP64 oracle.
Output format (oracle_pagerank_bucket_order): hand-computed examples.
"""
import numpy as np
import pytest

from belewitte_amd import graph
from oracle import binding as O

U = O.UNVISITED
D = np.float64(np.float32(0.85))


def _reference_depths(csr, root):
    """Independent shortest paths: networkx if present, else dense relaxation."""
    try:
        import networkx as nx
        G = nx.DiGraph()
        G.add_nodes_from(range(csr.V))
        for u in range(csr.V):
            for j in range(int(csr.off[u]), int(csr.off[u + 1])):
                G.add_edge(u, int(csr.col[j]))
        dist = nx.single_source_shortest_path_length(G, root)
        out = np.full(csr.V, U, dtype=np.int64)
        for v, d in dist.items():
            out[v] = d
        return out
    except ImportError:
        INF = 1 << 40
        d = np.full(csr.V, INF, dtype=np.int64)
        d[root] = 0
        src = np.repeat(np.arange(csr.V), np.diff(csr.off.astype(np.int64)))
        for _ in range(csr.V):
            nd = d.copy()
            np.minimum.at(nd, csr.col, d[src] + 1)
            if np.array_equal(nd, d):
                break
            d = nd
        d[d >= INF] = U
        return d


def _check_invariants(csr, root, d):
    d = d.astype(np.int64)
    assert d[root] == 0
    src = np.repeat(np.arange(csr.V), np.diff(csr.off.astype(np.int64)))
    dst = csr.col.astype(np.int64)
    reached = d != U
    # relaxation: d[v] <= d[u] + 1 for reached u
    ok = ~reached[src] | (d[dst] <= d[src] + 1)
    assert ok.all(), "edge relaxation violated"
    # every reached v != root has an in-neighbour at d[v] - 1
    has_parent = np.zeros(csr.V, dtype=bool)
    good = reached[src] & reached[dst] & (d[src] == d[dst] - 1)
    has_parent[dst[good]] = True
    need = reached.copy()
    need[root] = False
    assert has_parent[need].all(), "reached vertex without a parent one level up"
    # unreached vertices have no reached in-neighbour
    assert not (reached[src] & ~reached[dst]).any()


def _random_graphs():
    rng = np.random.default_rng(1234)
    for t in range(12):
        n = int(rng.integers(1, 60))
        m = int(rng.integers(0, 4 * n + 1))
        s = rng.integers(0, n, size=m)
        d = rng.integers(0, n, size=m)
        yield graph.from_edges(n, s, d, undirected=bool(t % 3 == 0))


@pytest.mark.parametrize("variant", range(8))
def test_bfs_variants_match_independent_shortest_paths(variant):
    for g in _random_graphs():
        csr = O.CSR.from_graph(g)
        for root in {0, g.vertex_count - 1}:
            ref = _reference_depths(csr, root)
            lv, counts = O.bfs_run(csr, root, [variant])
            assert np.array_equal(lv.astype(np.int64), ref)
            q, nlev = O.bfs_queue(csr, root)
            assert np.array_equal(q, lv)
            assert counts.size == nlev
            for k, c in enumerate(counts):
                assert c == int(np.sum(ref == k + 1))


def test_bfs_closed_forms():
    c = O.CSR.from_graph(graph.chain(40))
    d, n = O.bfs_queue(c, 0)
    assert d.tolist() == list(range(41)) and n == 41
    d, _ = O.bfs_queue(c, 40)  # sink: only itself
    assert d[40] == 0 and (d[:40] == U).all()
    s = O.CSR.from_graph(graph.star(25))
    d, n = O.bfs_queue(s, 0)
    assert d[0] == 0 and (d[1:] == 1).all() and n == 2
    H, W = 7, 9
    m = O.CSR.from_graph(graph.mesh(H, W))
    d, n = O.bfs_queue(m, 0)
    want = np.add.outer(np.arange(H), np.arange(W)).ravel()
    assert np.array_equal(d, want) and n == H + W - 1
    mu = O.CSR.from_graph(graph.mesh(H, W, undirected=True))
    d, _ = O.bfs_queue(mu, H * W - 1)
    want = np.add.outer(H - 1 - np.arange(H), W - 1 - np.arange(W)).ravel()
    assert np.array_equal(d, want)


def test_bfs_invariants_and_planted_mistakes():
    g = graph.rmat(12, 8, seed=77)
    csr = O.CSR.from_graph(g)
    root = g.max_out_degree_vertex()
    for variant in range(8):
        lv, _ = O.bfs_run(csr, root, [variant], warp=8, chunk=16)
        _check_invariants(csr, root, lv)
    good, _ = O.bfs_queue(csr, root)
    # Planted mistakes in the depths must be caught by the invariants.
    reached = np.nonzero((good != U) & (good > 0))[0]
    for bad in (lambda d: d.__setitem__(reached[5], d[reached[5]] + 1),
                lambda d: d.__setitem__(reached[7], d[reached[7]] - 1),
                lambda d: d.__setitem__(reached[-1], U)):
        d = good.copy()
        bad(d)
        with pytest.raises(AssertionError):
            _check_invariants(csr, root, d)


def test_bfs_root_out_of_range_and_schedules():
    g = graph.rmat(9, 4, seed=3)
    csr = O.CSR.from_graph(g)
    lv, counts = O.bfs_run(csr, g.vertex_count, [4])
    assert counts.size == 0
    assert O.bfs_queue(csr, g.vertex_count)[1] == 0
    want, _ = O.bfs_queue(csr, 1)
    rng = np.random.default_rng(5)
    for _ in range(5):
        lv, _ = O.bfs_run(csr, 1, list(rng.integers(0, 8, size=7)), warp=4, chunk=12)
        assert np.array_equal(lv, want)


def test_bfs_omp_port_matches():
    g = graph.rmat(13, 16, seed=9)
    csr = O.CSR.from_graph(g)
    root = g.max_out_degree_vertex()
    a, ca = O.bfs_run(csr, root, [4])
    b, cb = O.bfs_vertex_push_omp(csr, root, threads=4)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)


# ------------------------------------------------------------- PageRank

def _dense_pagerank(csr, iters, mode, quirk=False):
    """Power iteration on the dense column-stochastic matrix."""
    n = csr.V
    A = np.zeros((n, n))
    src = np.repeat(np.arange(n), np.diff(csr.off.astype(np.int64)))
    A[csr.col, src] = 1.0
    deg = A.sum(axis=0)
    P = np.divide(A, deg, out=np.zeros_like(A), where=deg > 0)
    x = np.full(n, 1.0 / n)
    for it in range(iters):
        s = (A @ x) if (quirk and it == 0) else (P @ x)
        dang = x[deg == 0].sum() if mode == "uniform" else 0.0
        x = (1 - D) / n + D * (s + dang / n)
    return x


@pytest.mark.parametrize("mode", ["drop", "uniform"])
def test_pagerank_f64_matches_dense_power_iteration(mode):
    for g in list(_random_graphs())[:8]:
        csr = O.CSR.from_graph(g)
        for quirk in (False, True):
            x = O.pagerank_f64(csr, 7, mode, nodiv_quirk=quirk)
            assert np.allclose(x, _dense_pagerank(csr, 7, mode, quirk), rtol=1e-12, atol=1e-15)


def test_pagerank_closed_forms():
    n = 50
    cyc = graph.from_edges(n, np.arange(n), (np.arange(n) + 1) % n)
    csr = O.CSR.from_graph(cyc)
    assert np.allclose(O.pagerank_f64(csr, 40, "drop"), 1.0 / n, rtol=0, atol=1e-15)
    # star 0 -> i: x0' = (1-d)/N; leaves (1-d)/N + d*x0/n; leaves dangle.
    k = 20
    st = O.CSR.from_graph(graph.star(k))
    N = k + 1
    x = O.pagerank_f64(st, 3, "drop")
    hub = (1 - D) / N
    assert x[0] == pytest.approx(hub, rel=1e-14)
    assert np.allclose(x[1:], (1 - D) / N + D * hub / k, rtol=1e-14)
    # uniform mode keeps the mass, drop mode loses the dangling part.
    xu = O.pagerank_f64(st, 12, "uniform")
    assert xu.sum() == pytest.approx(1.0, abs=1e-12)
    x1 = O.pagerank_f64(st, 1, "drop")
    x0 = np.full(N, 1.0 / N)
    lost = D * x0[1:].sum()
    assert x1.sum() == pytest.approx(1.0 - lost, abs=1e-12)


def test_pagerank_nodiv_quirk_first_iteration():
    g = graph.rmat(8, 8, seed=5)
    csr = O.CSR.from_graph(g)
    N = g.vertex_count
    indeg = np.diff(csr.roff.astype(np.int64))
    r, _ = O.pagerank_f32(csr, 8, 1, "drop")  # nodiv, 1 iteration, last -> undivided
    base = np.float32(np.float32(1.0) - np.float32(0.85)) / np.float32(N)
    want = (indeg / N) * D + float(base)
    assert np.allclose(r, want, rtol=1e-6)
    rd, _ = O.pagerank_f32(csr, 5, 1, "drop")  # div variant: divided sums
    assert not np.allclose(r, rd)


@pytest.mark.parametrize("mode", ["drop", "uniform"])
@pytest.mark.parametrize("variant", range(10))
def test_pagerank_f32_within_fp32_rounding_of_f64(variant, mode):
    g = graph.rmat(10, 8, seed=21)
    csr = O.CSR.from_graph(g)
    r, resid = O.pagerank_f32(csr, variant, 30, mode)
    quirk = variant in (8, 9) and mode == "drop"
    x = O.pagerank_f64(csr, 30, mode, nodiv_quirk=quirk)
    assert np.max(np.abs(r - x)) <= 2e-6 * x.max()
    assert resid.size == 30 and np.all(resid >= 0)
    if mode == "uniform":
        assert r.astype(np.float64).sum() == pytest.approx(1.0, abs=1e-5)


def test_pagerank_variants_agree_and_omp_port():
    g = graph.rmat(11, 16, seed=8)
    csr = O.CSR.from_graph(g)
    a, _ = O.pagerank_f32(csr, 8, 30, "uniform")
    b = O.pagerank_pull_omp(csr, 30, threads=4)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    ranks = [O.pagerank_f32(csr, v, 30, "uniform")[0] for v in range(10)]
    for r in ranks[1:]:
        assert np.max(np.abs(r - ranks[0])) < 1e-7


def test_bucket_order_hand_examples():
    assert O.bucket_order(np.array([3.0, 1.0, 2.0], np.float32)).tolist() == [1, 2, 0]
    # masking the low 16 mantissa bits: 1+2^-20 -> 1.0 (ties with id 0),
    # 1-2^-20 -> 0.99609375; rank order [2, 3, 0, 1], then the bucket
    # [0, n-1) is re-sorted by id and the last element stays: [0, 2, 3, 1].
    r = np.array([1.0, 1.0 + 2 ** -20, 0.5, 1.0 - 2 ** -20], np.float32)
    assert O.bucket_order(r).tolist() == [0, 2, 3, 1]
    r = np.array([5.0, 4.0, 3.0, 2.0, 1.0], np.float32)
    assert O.bucket_order(r).tolist() == [1, 2, 3, 4, 0]
    assert O.bucket_order(np.zeros(0, np.float32)).size == 0
