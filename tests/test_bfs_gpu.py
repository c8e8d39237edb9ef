"""BFS parity on the GPU: every reference implementation, bit-exact.

Depths must equal the oracle's textbook BFS and the per-launch frontier
counts must equal the oracle's restatement of the same kernel
(oracle/oracle.c, bfs/bfs.cpp:93-107 loop), for all 32 implementation names,
virtual-warp geometries, offset widths and random per-level schedules.
"""
import numpy as np
import pytest

from belewitte_amd import bfs, device, graph
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _graphs():
    yield graph.rmat(10, 8, seed=11)
    yield graph.rmat(12, 16, seed=12, undirected=True)
    yield graph.chain(300)
    yield graph.star(500)
    yield graph.mesh(17, 23, undirected=True)
    yield graph.erdos_renyi(3000, 6000, seed=5, undirected=False)
    yield graph.grid2d(40, 30, rewire=0.02, seed=4)


@pytest.fixture(scope="module")
def cases(gpu):
    out = []
    for g in _graphs():
        csr = O.CSR.from_graph(g)
        dg = device.DeviceGraph(g, reps=device.REPRESENTATIONS)
        roots = sorted({g.max_out_degree_vertex(), 0, g.vertex_count // 2})
        out.append((g, csr, dg, roots))
    yield out
    for _, _, dg, _ in out:
        dg.free()


@pytest.mark.parametrize("name", sorted(bfs.IMPLEMENTATIONS))
def test_all_implementations_bit_exact(cases, name):
    k, _ = bfs.IMPLEMENTATIONS[name]
    for g, csr, dg, roots in cases:
        for root in roots:
            print(f"case {g.name} root={root} impl={name}", flush=True)
            res = bfs.run(dg, root, name)
            want, nlev = O.bfs_queue(csr, root)
            assert np.array_equal(res.levels, want), (g.name, root, name)
            _, counts = O.bfs_run(csr, root, [k])
            assert res.num_levels == nlev
            assert np.array_equal(res.counts, counts), (g.name, root, name)


@pytest.mark.parametrize("warp,chunk", [(1, 1), (2, 6), (8, 8), (16, 48), (32, 32), (64, 64), (64, 128)])
@pytest.mark.parametrize("name", ["vertex-push-warp", "vertex-pull-warp-blockreduce"])
def test_warp_geometries(cases, name, warp, chunk):
    for g, csr, dg, roots in cases[:3]:
        root = roots[0]
        res = bfs.run(dg, root, name, warp=warp, chunk=chunk)
        want, _ = O.bfs_queue(csr, root)
        assert np.array_equal(res.levels, want), (g.name, name, warp, chunk)


@pytest.mark.parametrize("block", [64, 128, 1024])
def test_block_sizes(cases, block):
    g, csr, dg, roots = cases[0]
    for name in ("edge-list-blockreduce", "vertex-pull-warpreduce", "vertex-push-warp-bulk"):
        res = bfs.run(dg, roots[0], name, block=block)
        want, _ = O.bfs_queue(csr, roots[0])
        assert np.array_equal(res.levels, want)


def test_random_schedules_bit_exact(cases):
    rng = np.random.default_rng(7)
    names = sorted(bfs.IMPLEMENTATIONS)
    for g, csr, dg, roots in cases:
        want, nlev = O.bfs_queue(csr, roots[0])
        for trial in range(4):
            sched = [names[i] for i in rng.integers(0, len(names), size=max(nlev, 1))]
            res = bfs.run(dg, roots[0], schedule=sched)
            assert np.array_equal(res.levels, want), (g.name, sched)
            assert res.num_levels == nlev


def test_uint64_offsets(gpu):
    g = graph.rmat(11, 8, seed=21)
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g, offset_bytes=8)
    root = g.max_out_degree_vertex()
    want, _ = O.bfs_queue(csr, root)
    for name in ("vertex-push", "vertex-pull-bulk", "vertex-push-warp-warpreduce", "vertex-pull-warp"):
        res = bfs.run(dg, root, name)
        assert np.array_equal(res.levels, want), name
    dg.free()


def test_edge_cases(gpu):
    # root >= V: empty output (bfs/bfs.cpp:53)
    g = graph.chain(10)
    dg = device.DeviceGraph(g)
    res = bfs.run(dg, 11, "vertex-push")
    assert res.levels.size == 0 and res.num_levels == 0
    # isolated root: one launch that finds nothing
    res = bfs.run(dg, 10, "vertex-push")
    assert res.num_levels == 1 and res.levels[10] == 0
    assert np.all(res.levels[:10] == bfs.UNVISITED)
    dg.free()
    # graph without edges
    e = graph.from_edges(5, [], [])
    de = device.DeviceGraph(e, reps=device.REPRESENTATIONS)
    for name in ("edge-list", "vertex-pull-warp", "struct-edge-list-bulk"):
        res = bfs.run(de, 2, name)
        assert res.levels.tolist() == [bfs.UNVISITED] * 2 + [0] + [bfs.UNVISITED] * 2
    de.free()


def test_cfg1_vertex_push_rmat_s16(gpu):
    """R-MAT scale 16, edge factor 16, uint64 offsets, max-out-degree root:
    vertex-push in all four reductions, bit-exact."""
    g = graph.rmat(16, 16, seed=1)
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g, offset_bytes=8)
    root = g.max_out_degree_vertex()
    want, nlev = O.bfs_queue(csr, root)
    _, counts = O.bfs_run(csr, root, [4])
    for suffix in bfs.REDUCTIONS:
        res = bfs.run(dg, root, "vertex-push" + suffix)
        assert np.array_equal(res.levels, want)
        assert np.array_equal(res.counts, counts)
    dg.free()
