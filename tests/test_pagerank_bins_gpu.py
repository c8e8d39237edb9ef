"""The binned PageRank iteration (propagation blocking) on the GPU.

Ranks against the oracle: the FP32 restatement of the reference's
vertex-pull-nodiv (pagerank/vertex_pull_nodiv.cu:3-29 with
consolidateRankNoDiv, pagerank/pagerank.cu:142-165) and the FP64
recurrence.  Bar: 1e-6 max-abs against the FP32 oracle, widened by the
FP32 oracle's own error against FP64 (its sequential sums drift on hub
rows), and 1e-6 against the FP64 recurrence.  The row accumulators are
64-bit fixed point (integer sums, independent of the order of the
atomics), so the ranks repeat bit for bit from run to run.
"""
import numpy as np
import pytest

from belewitte_amd import device, graph, pagerank
from oracle import binding as O
from test_degenerate_gpu import degenerate_graphs

pytestmark = pytest.mark.gpu

TOL = 1e-6
SETTINGS = [(13, 1, 8), (13, 64, 16), (14, 3, 8), (14, 64, 16)]


def _hub_graph():
    r = graph.rmat(15, 4, seed=41)
    s, d = r.edges()
    hub_src = np.arange(1, 20001, dtype=np.uint32)
    return graph.from_edges(r.vertex_count, np.concatenate([s, hub_src]),
                            np.concatenate([d, np.zeros(hub_src.size, dtype=np.uint32)]))


@pytest.fixture(scope="module")
def cases(gpu):
    out = []
    for g in (graph.rmat(12, 16, seed=31), graph.erdos_renyi(2000, 9000, seed=3, undirected=False),
              graph.rmat(11, 8, seed=32, undirected=True), graph.star(300), graph.chain(50), _hub_graph()):
        out.append((g, O.CSR.from_graph(g), device.DeviceGraph(g)))
    yield out
    for _, _, dg in out:
        dg.free()


def _check(g, csr, ranks, resid, mode, quirk, iters):
    """Within 1e-6 of the FP32 oracle widened by its own error (the
    reference's sequential FP32 sums; in DROP mode without the undivided
    first iteration its vertex-pull variant, which is the same
    recurrence), within 1e-6 of the FP64 recurrence, and never less
    accurate than the FP32 oracle."""
    x = O.pagerank_f64(csr, iters, mode, nodiv_quirk=quirk)
    variant = "vertex-pull-nodiv" if (mode == "uniform" or quirk) else "vertex-pull"
    want, _ = O.pagerank_f32(csr, pagerank.variant_id(variant), iters, mode)
    ref_err = float(np.max(np.abs(want.astype(np.float64) - x)))
    err = float(np.max(np.abs(ranks.astype(np.float64) - want)))
    assert err <= TOL + ref_err, (g.name, mode, quirk, err, ref_err)
    err64 = float(np.max(np.abs(ranks - x)))
    assert err64 <= TOL and err64 <= ref_err + 1e-7, (g.name, mode, quirk, err64, ref_err)
    assert np.all(np.isfinite(resid))


@pytest.mark.parametrize("rows_log2,chunk_tasks,lanes", SETTINGS)
@pytest.mark.parametrize("mode,quirk,iters", [("uniform", False, 30), ("drop", True, 100), ("drop", False, 20)])
def test_binned_matches_oracle(cases, rows_log2, chunk_tasks, lanes, mode, quirk, iters):
    for g, csr, dg in cases:
        print(f"case {g.name} binned {rows_log2}/{chunk_tasks}/{lanes} mode={mode} quirk={quirk}", flush=True)
        fp = pagerank.FusedPageRank(dg, max_iterations=iters)
        pull, rp, _ = fp.run(iters, mode, nodiv_quirk=quirk)
        fp.bin(rows_log2, chunk_tasks, lanes)
        ranks, resid, _ = fp.run(iters, mode, nodiv_quirk=quirk)
        _check(g, csr, ranks, resid, mode, quirk, iters)
        # The residuals follow the pull kernel's (Kahan-compensated) ones to
        # FP32 rounding.
        assert np.allclose(resid, rp, rtol=1e-3, atol=1e-6), (g.name, resid[:4], rp[:4])
        again, ragain, _ = fp.run(iters, mode, nodiv_quirk=quirk)
        assert np.array_equal(again.view(np.uint32), ranks.view(np.uint32)), g.name
        assert np.array_equal(ragain, resid), g.name
        # Back to the pull kernels: the same ranks as before, bit for bit.
        fp.bin(0)
        back, _, _ = fp.run(iters, mode, nodiv_quirk=quirk)
        assert np.array_equal(back.view(np.uint32), pull.view(np.uint32))
        fp.close()


@pytest.mark.parametrize("rows_log2,chunk_tasks,lanes", [(13, 2, 16), (13, 1, 8), (14, 64, 8)])
def test_binned_degree_ordered(gpu, rows_log2, chunk_tasks, lanes):
    """Several partitions, many chunks each (R-MAT scale 16 relabelled by
    decreasing out-degree: partition 0 holds the hubs)."""
    g = graph.rmat(16, 16, seed=51)
    csr = O.CSR.from_graph(g)
    fp = pagerank.FusedPageRank.for_graph(g, max_iterations=30, reorder=True)
    fp.tune(0, 0, 1, 2, 16)
    pull, _, _ = fp.run(30, "uniform")
    fp.bin(rows_log2, chunk_tasks, lanes)
    info = fp.bin_info()
    assert info["chunks"] > 8 and info["blocks"] >= 16 and 0 < info["pairs"] < g.edge_count
    ranks, resid, _ = fp.run(30, "uniform")
    _check(g, csr, ranks, resid, "uniform", False, 30)
    assert float(np.max(np.abs(ranks - pull))) <= 2 * TOL
    fp.close()


@pytest.mark.parametrize("mode", ["drop", "uniform"])
def test_binned_degenerate(gpu, mode):
    for g in degenerate_graphs():
        csr = O.CSR.from_graph(g)
        dg = device.DeviceGraph(g)
        fp = pagerank.FusedPageRank(dg, max_iterations=10)
        fp.bin(13, 1, 8)
        ranks, resid, _ = fp.run(10, mode)
        x = O.pagerank_f64(csr, 10, mode)
        assert float(np.max(np.abs(ranks - x))) <= TOL, g.name
        assert np.all(np.isfinite(resid))
        fp.close()
        dg.free()


@pytest.mark.parametrize("mode,quirk", [("drop", True), ("drop", False), ("uniform", False)])
def test_binned_dense_multigraph(gpu, tmp_path, mode, quirk):
    """16 vertices and 60,000 edges, most of them repeated into vertex 0 (a
    .graph file keeps repeated edges): after the undivided first iteration
    of the nodiv quirk that row's sum is about 3,200, far above the rank
    mass of 1, which the fixed-point row accumulators must hold without
    overflow."""
    from graphfiles import write_v1
    rng = np.random.default_rng(12)
    src = rng.integers(0, 16, 60000)
    dst = np.where(rng.random(60000) < 0.85, 0, rng.integers(0, 16, 60000))
    write_v1(tmp_path / "multi.graph", 16, src, dst)
    g = graph.load(tmp_path / "multi.graph")
    assert g.edge_count == 60000
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g)
    fp = pagerank.FusedPageRank(dg, max_iterations=12)
    fp.bin(13, 1, 8)
    for iters in (1, 2, 12):
        ranks, resid, _ = fp.run(iters, mode, nodiv_quirk=quirk)
        x = O.pagerank_f64(csr, iters, mode, nodiv_quirk=quirk)
        assert float(np.max(np.abs(ranks - x) / x)) <= 1e-6, (mode, quirk, iters, ranks[:4], x[:4])
        assert np.all(np.isfinite(resid))
    fp.close()
    dg.free()


def test_binned_rejects_bad_settings(gpu):
    g = graph.rmat(10, 8, seed=1)
    dg = device.DeviceGraph(g)
    fp = pagerank.FusedPageRank(dg, max_iterations=5)
    for args in ((12, 1, 8), (15, 1, 8), (13, 0, 8), (13, 1, 12)):
        with pytest.raises(Exception):
            fp.bin(*args)
    fp.close()
    dg.free()


def test_binned_failed_shape_keeps_plan(gpu):
    """A chain of 2^24 + 2 vertices: 1025 partitions at 2^14 rows, 2049 at
    2^13, one more than the layout holds.  The refused 2^13 shape leaves
    the plan binned at 2^14, and its ranks still match the oracle."""
    g = graph.chain((1 << 24) + 1)
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g)
    fp = pagerank.FusedPageRank(dg, max_iterations=5)
    fp.bin(14, 64, 8)
    before = fp.bin_info()
    with pytest.raises(Exception, match="too many partitions"):
        fp.bin(13, 64, 8)
    assert fp.binning == (14, 64, 8)
    assert fp.bin_info() == before and before["pairs"] > 0
    ranks, resid, _ = fp.run(5, "uniform")
    assert float(np.max(np.abs(ranks - O.pagerank_f64(csr, 5, "uniform")))) <= TOL
    assert np.all(np.isfinite(resid))
    fp.close()
    dg.free()


@pytest.mark.parametrize("shape", [(13, 64, 8), (14, 128, 8), (13, 32, 16)])
def test_binned_agrees_with_pull_at_scale_20(gpu, shape):
    """R-MAT scale 20, edge factor 16, degree-ordered: the binned iteration
    against the Kahan-compensated pull kernel (same recurrence) to FP32
    rounding of the ranks, and UNIFORM mode conserving the rank mass."""
    g = graph.rmat(20, 16, seed=7)
    fp = pagerank.FusedPageRank.for_graph(g, max_iterations=30, reorder=True)
    fp.tune(0, 0, 1, 2, 16)
    pull, rp, _ = fp.run(30, "uniform")
    fp.bin(*shape)
    ranks, resid, _ = fp.run(30, "uniform")
    assert float(np.max(np.abs(ranks - pull))) <= 4e-7 * float(pull.max()), shape
    assert abs(float(np.sum(ranks, dtype=np.float64)) - 1.0) <= 1e-5
    assert np.allclose(resid, rp, rtol=1e-3, atol=1e-6)
    fp.close()


def _plan_graphs():
    yield graph.rmat(16, 16, seed=51)
    yield graph.rmat(17, 8, seed=52, undirected=True)
    g, _ = graph.rmat(16, 32, seed=53).reorder_by_degree()
    yield g
    yield _hub_graph()
    yield graph.erdos_renyi(30000, 200000, seed=54, undirected=False)
    yield graph.grid2d(300, 200, rewire=0.05, seed=55)
    yield graph.star(70000)
    yield graph.chain(40000)
    yield from degenerate_graphs()


@pytest.mark.parametrize("shape", [(13, 1, 8), (13, 64, 8), (14, 3, 16), (14, 128, 8)])
def test_device_plan_equals_host_plan(gpu, shape):
    """The plan built on the device (sorts, scans and the greedy block cut
    on the GPU) is the host builder's, array for array and bit for bit."""
    for g in _plan_graphs():
        dg = device.DeviceGraph(g)
        fp = pagerank.FusedPageRank(dg, max_iterations=2)
        try:
            print(f"plan {g.name} {shape}", flush=True)
            fp.bin(*shape)
            dev = {k: fp.bin_array(k) for k in pagerank.FusedPageRank.BIN_ARRAYS}
            info_dev = fp.bin_info()
            fp.bin(*shape, on_host=True)
            host = {k: fp.bin_array(k) for k in pagerank.FusedPageRank.BIN_ARRAYS}
            assert fp.bin_info() == info_dev, g.name
            for k in dev:
                assert dev[k].shape == host[k].shape, (g.name, shape, k, dev[k].shape, host[k].shape)
                assert np.array_equal(dev[k], host[k]), (g.name, shape, k)
        finally:
            fp.close()
            dg.free()


def test_device_plan_ranks_at_scale_20(gpu):
    """The device-built plan at R-MAT scale 20 (degree-ordered) against the
    oracle, both modes."""
    g0 = graph.rmat(20, 16, seed=56)
    g, perm = g0.reorder_by_degree()
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g)
    fp = pagerank.FusedPageRank(dg, max_iterations=30)
    try:
        for shape in ((14, 128, 8), (13, 64, 8)):
            fp.bin(*shape)
            for mode, quirk, iters in (("uniform", False, 30), ("drop", True, 20)):
                ranks, resid, _ = fp.run(iters, mode, nodiv_quirk=quirk)
                _check(g, csr, ranks, resid, mode, quirk, iters)
    finally:
        fp.close()
        dg.free()


def test_device_relabel_equals_host_relabel(gpu):
    """bw_graph_relabel_by_degree gives the graph Graph.reorder_by_degree()
    gives on the host: the same permutation, offsets and sorted column
    lists in both stored directions; directed and undirected graphs, ties
    and vertices without edges."""
    graphs = [graph.rmat(16, 16, seed=61), graph.rmat(15, 8, seed=62, undirected=True), _hub_graph(),
              graph.erdos_renyi(20000, 90000, seed=63, undirected=False), graph.star(5000), graph.chain(3000),
              graph.grid2d(100, 70, rewire=0.02, seed=64)]
    graphs += list(degenerate_graphs())
    for g in graphs:
        want, perm_want = g.reorder_by_degree()
        dg = device.DeviceGraph(g)
        rg, perm = dg.relabel_by_degree()
        try:
            assert np.array_equal(perm, perm_want), g.name
            for rev, ko, kc in ((False, "off", "col"), (True, "roff", "rcol")):
                off = rg.array(ko).download()
                assert np.array_equal(off, want.offsets32(rev)), (g.name, rev)
                if g.edge_count:
                    cols = rg.array(kc).download()[:g.edge_count]
                    assert np.array_equal(cols, want.cols(rev)), (g.name, rev)
        finally:
            rg.free()
            dg.free()


def test_device_relabel_pipeline_ranks(gpu):
    """Relabel and plan both on the device, from the generated order: ranks
    (returned in the original order) against the oracle."""
    g = graph.rmat(18, 16, seed=65)
    csr = O.CSR.from_graph(g)
    fp = pagerank.FusedPageRank.for_graph(g, max_iterations=20, reorder=True, on_device=True)
    try:
        fp.bin(14, 128, 8)
        for mode, quirk, iters in (("uniform", False, 20), ("drop", True, 20)):
            ranks, resid, _ = fp.run(iters, mode, nodiv_quirk=quirk)
            _check(g, csr, ranks, resid, mode, quirk, iters)
    finally:
        fp.close()


@pytest.mark.parametrize("rows_log2,lanes", [(13, 8), (14, 8), (13, 16), (14, 16)])
@pytest.mark.parametrize("quirk", [False, True])
def test_binned_drop_ranks_bit_exact(cases, rows_log2, lanes, quirk):
    """With every partition in one chunk (4096 tasks of 2048 edges), a
    DROP-mode run's ranks are fixed by the fixed-point arithmetic alone:
    they equal oracle_pagerank_binned_fixed bit for bit, so a single edge
    added twice or missed anywhere would show."""
    for g, csr, dg in cases:
        print(f"case {g.name} bit-exact binned {rows_log2}/{lanes} quirk={quirk}", flush=True)
        fp = pagerank.FusedPageRank(dg, max_iterations=20)
        fp.bin(rows_log2, 4096, lanes)
        parts = -(-g.vertex_count // (1 << rows_log2))
        assert fp.bin_info()["chunks"] <= parts
        for iters in (1, 2, 20):
            ranks, _, _ = fp.run(iters, "drop", nodiv_quirk=quirk)
            want = O.pagerank_binned_fixed(csr, iters, nodiv_quirk=quirk)
            bad = np.flatnonzero(ranks.view(np.uint32) != want.view(np.uint32))
            assert bad.size == 0, (g.name, iters, bad[:5], ranks[bad[:5]], want[bad[:5]])
        fp.close()


@pytest.mark.parametrize("quirk", [False, True])
def test_binned_drop_ranks_bit_exact_scale_20(gpu, quirk):
    """R-MAT scale 20 (edge factor 16) relabelled by degree on the device,
    2^13-row partitions of one chunk each (the hub partition holds 7.4 M of
    the 16.1 M edges): bit for bit against the oracle's fixed-point
    restatement after 1 and 10 iterations."""
    g = graph.rmat(20, 16, seed=7)
    fp = pagerank.FusedPageRank.for_graph(g, max_iterations=10, reorder=True, on_device=True)
    try:
        fp.bin(13, 4096, 8)
        assert fp.bin_info()["chunks"] <= g.vertex_count >> 13
        rg, _ = g.reorder_by_degree()
        csr = O.CSR.from_graph(rg)
        for iters in (1, 10):
            ranks, _, _ = fp.run(iters, "drop", nodiv_quirk=quirk)  # in the original order
            want = O.pagerank_binned_fixed(csr, iters, nodiv_quirk=quirk)[fp.perm]
            bad = np.flatnonzero(ranks.view(np.uint32) != want.view(np.uint32))
            assert bad.size == 0, (iters, bad.size, bad[:5], ranks[bad[:5]], want[bad[:5]])
    finally:
        fp.close()


def test_autotuned_shape_matches_oracle(gpu):
    """The autotuner times every candidate shape (binned and pull) on
    R-MAT scale 18 in both vertex orders and keeps one; the ranks of the
    kept shape match the oracle."""
    g = graph.rmat(18, 16, seed=67)
    csr = O.CSR.from_graph(g)
    for reorder in (True, False):
        fp = pagerank.FusedPageRank.for_graph(g, max_iterations=20, reorder=reorder, on_device=True)
        try:
            table = fp.autotune()
            print(f"autotune reorder={reorder}: {table} -> {fp.kernel_names()}", flush=True)
            assert len(table) >= 6
            for mode, quirk, iters in (("uniform", False, 20), ("drop", True, 20)):
                ranks, resid, _ = fp.run(iters, mode, nodiv_quirk=quirk)
                _check(g, csr, ranks, resid, mode, quirk, iters)
        finally:
            fp.close()
