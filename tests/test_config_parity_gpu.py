"""Oracle parity at the configurations' full sizes.

The reference accepts a timing run only when its output hash equals the
reference implementation's on the same input
(benchmark-analysis/ingest-src/Jobs.hs:230-241).  Here every configuration
is checked against the oracle at the size its figures are measured on:

* cfg2: PageRank on the directed R-MAT scale 22 (edge factor 16, seed 2):
  the reference's 10 variants and the fused iteration (binned on the
  degree-ordered and on the generated order, and the pull kernels), DROP
  and UNIFORM modes, 30 iterations.
* cfg3: BFS on the Erdos-Renyi graph (2^24 vertices, 2^27 undirected pairs,
  seed 3) and on the 4096 x 4096 grid with 1% of its edges rewired
  (seed 4), 16 sources each (seed 10): the learned switch, Beamer's rule,
  every engine variant, the exhaustive per-level collector and the
  reference's kernel families through the reference level loop.
* cfg4: 64 sources on the directed R-MAT scale 20 (seed 6) in one
  multi-source search, every source checked; 256 R-MAT graphs searched as
  one disjoint union, every component checked against its own graph; both
  again after the autotuner has chosen their thresholds and queue sizes.
* cfg5: the learned switch on the undirected R-MAT scale 25 (seed 5) from
  16 sources (seed 8), with a fixed native, a fixed reference-family and
  the Beamer run beside it.

The depth oracle is the textbook queue BFS (oracle_bfs_queue), run for
several sources at once on the host cores; per-level counts follow from
its depths (the count of launch k is the number of vertices at depth
k + 1, the last launch finds none).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from belewitte_amd import batch, bfs, bfs_native, device, graph, pagerank, switch
from oracle import binding as O

pytestmark = pytest.mark.gpu

TOL = 1e-6
UNVISITED = O.UNVISITED
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def oracle_depths(csr, roots):
    """Queue-BFS depths from each root, THREADS searches at a time (the
    oracle call releases the interpreter lock)."""
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        return [d for d, _ in ex.map(lambda r: O.bfs_queue(csr, int(r)), roots)]


def level_counts(depth):
    """Per-launch frontier counts of the reference loop (bfs/bfs.cpp:93-107)
    implied by exact depths."""
    reached = depth[depth != UNVISITED]
    hist = np.bincount(reached.astype(np.int64), minlength=1) if reached.size else np.zeros(1, np.int64)
    return np.append(hist[1:], 0).astype(np.uint32)


# --------------------------------------------------------------------- cfg2

@pytest.fixture(scope="module")
def cfg2(gpu):
    g = graph.rmat(22, 16, seed=2)
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g, reps=device.REPRESENTATIONS)
    ref = {}
    for mode in ("drop", "uniform"):
        for name in ("vertex-pull", "vertex-pull-nodiv"):
            ref[mode, name] = O.pagerank_f32(csr, pagerank.variant_id(name), 30, mode)[0]
        ref[mode, "f64"] = O.pagerank_f64(csr, 30, mode)
    ref["drop", "f64-quirk"] = O.pagerank_f64(csr, 30, "drop", nodiv_quirk=True)
    yield g, csr, dg, ref
    dg.free()


def _pr_bars(ref, mode, nodiv):
    """The FP32 oracle of the pull variant with the same division form and
    start, and the FP64 recurrence it follows."""
    want = ref[mode, "vertex-pull-nodiv" if nodiv else "vertex-pull"]
    x = ref[mode, "f64-quirk"] if (nodiv and mode == "drop") else ref[mode, "f64"]
    return want, x, float(np.max(np.abs(want.astype(np.float64) - x)))


@pytest.mark.parametrize("mode", ["drop", "uniform"])
@pytest.mark.parametrize("variant", pagerank.VARIANTS)
def test_cfg2_pagerank_variant_at_scale_22(cfg2, variant, mode):
    """Deterministic pull variants bit for bit against the oracle; the
    others (atomic sums, whose order is not fixed in the reference either)
    within 1e-6 of it, and within 1e-6 plus the oracle's own error of the
    FP64 recurrence."""
    g, csr, dg, ref = cfg2
    nodiv = pagerank.variant_id(variant) in pagerank.NODIV
    print(f"cfg2 {g.name} variant={variant} mode={mode}", flush=True)
    ranks, resid = pagerank.run(dg, variant, iterations=30, mode=mode)
    want, x, ref_err = _pr_bars(ref, mode, nodiv)
    err = float(np.max(np.abs(ranks.astype(np.float64) - want)))
    err64 = float(np.max(np.abs(ranks - x)))
    print(f"cfg2 {variant} {mode}: max-abs vs FP32 oracle {err:.3e}, vs FP64 {err64:.3e} "
          f"(oracle vs FP64 {ref_err:.3e})", flush=True)
    if variant in ("vertex-pull", "vertex-pull-nodiv"):
        assert np.array_equal(ranks.view(np.uint32), want.view(np.uint32)), (variant, mode, err)
    assert err <= TOL, (variant, mode, err)
    assert err64 <= TOL + ref_err, (variant, mode, err64, ref_err)
    assert np.all(np.isfinite(resid))


@pytest.mark.parametrize("mode", ["drop", "uniform"])
def test_cfg2_fused_at_scale_22(cfg2, mode):
    """The fused iteration in each form on the same graph: binned on the
    degree-ordered upload (the benchmark's kernels) and on the generated
    order, and the pull kernels; 1e-6 of the FP32 oracle widened by the
    oracle's own error, and 1e-6 of the FP64 recurrence."""
    g, csr, dg, ref = cfg2
    quirk = mode == "drop"
    want, x, ref_err = _pr_bars(ref, mode, True)
    forms = []
    fp = pagerank.FusedPageRank.for_graph(g, max_iterations=30, reorder=True)
    forms.append(("binned degree-ordered", fp, None))
    nat = pagerank.FusedPageRank(dg, max_iterations=30)
    forms.append(("binned generated order", nat, (14, 128, 8)))
    # the shape the autotuner picks for the generated order (its bench figure)
    forms.append(("binned generated order, 80-task chunks", nat, (14, 80, 8)))
    forms.append(("pull kernels", nat, 0))
    for label, f, shape in forms:
        if shape is None:
            f.bin(14, 128, 8)
        elif shape == 0:
            f.bin(0)
            f.tune(0, 768, 1, 2, 16)
        else:
            f.bin(*shape)
        ranks, resid, _ = f.run(30, mode, nodiv_quirk=quirk)
        err = float(np.max(np.abs(ranks.astype(np.float64) - want)))
        err64 = float(np.max(np.abs(ranks - x)))
        print(f"cfg2 fused {label} {mode}: vs FP32 oracle {err:.3e}, vs FP64 {err64:.3e}", flush=True)
        assert err <= TOL + ref_err, (label, mode, err, ref_err)
        assert err64 <= TOL, (label, mode, err64)
        assert np.all(np.isfinite(resid))
    fp.close()
    nat.close()


# --------------------------------------------------------------------- cfg3

def _cfg3_graph(which):
    if which == "erdos-renyi":
        return graph.erdos_renyi(1 << 24, 134217728, seed=3, undirected=True)
    return graph.grid2d(4096, 4096, rewire=0.01, seed=4)


@pytest.mark.parametrize("which", ["erdos-renyi", "grid"])
def test_cfg3_bfs_full_size(gpu, which):
    g = _cfg3_graph(which)
    csr = O.CSR.from_graph(g)
    src = [int(s) for s in graph.random_sources(g, 16, seed=10)]
    want = oracle_depths(csr, src)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    try:
        for s, w in zip(src, want):
            print(f"cfg3 {g.name} source {s} learned", flush=True)
            r = eng.run(s, "learned", "td")
            assert np.array_equal(r.levels, w), (g.name, s)
            assert np.array_equal(r.counts, level_counts(w)), (g.name, s)
        print(f"cfg3 {g.name} source {src[0]} beamer", flush=True)
        assert np.array_equal(eng.run(src[0], "beamer", "td").levels, want[0])
        for k, v in enumerate(bfs_native.VARIANTS):
            s, w = src[k], want[k]
            print(f"cfg3 {g.name} source {s} fixed {v}", flush=True)
            r = eng.run(s, "fixed", v)
            assert np.array_equal(r.levels, w), (g.name, s, v)
        # The exhaustive collector (every variant timed at every level from
        # one saved state) ends on the oracle's depths.
        s, w = src[-1], want[-1]
        print(f"cfg3 {g.name} source {s} exhaustive", flush=True)
        lv, counts, recs = eng.exhaustive(s)
        assert np.array_equal(lv, w) and np.array_equal(counts, level_counts(w))
        assert len(recs) == counts.size
    finally:
        eng.close()
    # The reference's kernel families through the reference level loop.
    s, w = src[0], want[0]
    for k, name in enumerate(bfs.KERNELS):
        dg.ensure(bfs.REPRESENTATION[k])
        print(f"cfg3 {g.name} source {s} reference {name}", flush=True)
        r = bfs.run(dg, s, name + "-warpreduce", time_levels=False)
        assert np.array_equal(r.levels, w), (g.name, name)
        assert np.array_equal(r.counts, level_counts(w)), (g.name, name)
    dg.free()


# --------------------------------------------------------------------- cfg4

def test_cfg4_multi_source_every_source(gpu):
    g = graph.rmat(20, 16, seed=6)
    csr = O.CSR.from_graph(g)
    src = [int(s) for s in graph.random_sources(g, 64, seed=9)]
    want = oracle_depths(csr, src)
    dg = device.DeviceGraph(g)
    ms = bfs_native.MultiSourceBFS(dg)
    try:
        for tuning in ((32, 4, 128), (64, 8, 32), (32, 0, 128)):
            ms.set_tuning(*tuning)
            print(f"cfg4 {g.name} 64 sources tuning {tuning}", flush=True)
            depths, _, _ = ms.run(src)
            for i, w in enumerate(want):
                assert np.array_equal(depths[i], w), (tuning, i, src[i])
    finally:
        ms.close()
        dg.free()


def test_cfg4_autotuned_every_source_and_component(gpu):
    """The autotuner at the batch sizes: the 64-source search on R-MAT
    scale 20 and the 256-graph union, each tuned (thresholds, then queue
    sizes) and rerun with the chosen parameters; every source and every
    component checked against the oracle."""
    g = graph.rmat(20, 16, seed=6)
    csr = O.CSR.from_graph(g)
    src = [int(s) for s in graph.random_sources(g, 64, seed=9)]
    want = oracle_depths(csr, src)
    dg = device.DeviceGraph(g)
    ms = bfs_native.MultiSourceBFS(dg)
    try:
        tuned = batch.autotune_msbfs(ms, src)
        print(f"cfg4 64 sources tuned {tuned.params}", flush=True)
        depths, _, _ = ms.run(src)
        for i, w in enumerate(want):
            assert np.array_equal(depths[i], w), (tuned.params, i, src[i])
    finally:
        ms.close()
        dg.free()

    graphs = batch.small_rmat_batch(256, seed=7)
    u, base = batch.union(graphs)
    roots = [base[i] + gi.max_out_degree_vertex() for i, gi in enumerate(graphs)]
    csrs = [O.CSR.from_graph(gi) for gi in graphs]
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        want = list(ex.map(lambda i: O.bfs_queue(csrs[i], graphs[i].max_out_degree_vertex())[0],
                           range(len(graphs))))
    dg = device.DeviceGraph(u)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    try:
        tuned = batch.autotune_engine(eng, roots, "learned", "td")
        assert (eng.tuning + eng.queues) == tuned.params
        print(f"cfg4 union tuned {tuned.params}", flush=True)
        runs = {"learned": eng.run(roots, "learned", "td"), "td": eng.run(roots, "fixed", "td"),
                "edge-list": eng.run(roots, "fixed", "edge-list")}
    finally:
        eng.close()
        dg.free()
    for label, r in runs.items():
        for i, gi in enumerate(graphs):
            assert np.array_equal(r.levels[base[i]:base[i] + gi.vertex_count], want[i]), (label, i, gi.name)


def test_cfg4_union_every_component(gpu):
    graphs = batch.small_rmat_batch(256, seed=7)
    u, base = batch.union(graphs)
    roots = [base[i] + gi.max_out_degree_vertex() for i, gi in enumerate(graphs)]
    dg = device.DeviceGraph(u)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    try:
        runs = {"learned": eng.run(roots, "learned", "td"), "beamer": eng.run(roots, "beamer", "td")}
        for v in bfs_native.VARIANTS:
            runs[v] = eng.run(roots, "fixed", v)
    finally:
        eng.close()
        dg.free()
    csrs = [O.CSR.from_graph(gi) for gi in graphs]
    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        want = list(ex.map(lambda i: O.bfs_queue(csrs[i], graphs[i].max_out_degree_vertex())[0],
                           range(len(graphs))))
    for label, r in runs.items():
        for i, gi in enumerate(graphs):
            got = r.levels[base[i]:base[i] + gi.vertex_count]
            assert np.array_equal(got, want[i]), (label, i, gi.name)
        assert r.levels.size == u.vertex_count


# --------------------------------------------------------------------- cfg5

def test_cfg5_learned_switch_scale_25(gpu):
    g = graph.rmat(25, 16, seed=5, undirected=True)
    csr = O.CSR.from_graph(g)
    src = [int(s) for s in graph.random_sources(g, 16, seed=8)]
    want = oracle_depths(csr, src)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    try:
        for s, w in zip(src, want):
            print(f"cfg5 {g.name} source {s} learned", flush=True)
            r = eng.run(s, "learned", "td")
            assert np.array_equal(r.levels, w), s
            assert np.array_equal(r.counts, level_counts(w)), s
        for v, s, w in (("td", src[0], want[0]), ("bu-wave", src[1], want[1]), ("edge-list", src[2], want[2])):
            print(f"cfg5 {g.name} source {s} fixed {v}", flush=True)
            assert np.array_equal(eng.run(s, "fixed", v).levels, w), (v, s)
        print(f"cfg5 {g.name} source {src[3]} beamer", flush=True)
        assert np.array_equal(eng.run(src[3], "beamer", "td").levels, want[3])
    finally:
        eng.close()
        dg.free()
