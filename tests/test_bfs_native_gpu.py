"""The device-driven BFS engine, bit-exact against the oracle.

Depths must equal the oracle's textbook queue BFS and the per-level counts
the oracle's restatement of the reference level loop (bfs/bfs.cpp:93-107:
newly discovered vertices per level, last 0), for every native variant,
fixed and per-level schedules, host-stepped and hipGraph-batched loops,
the learned policy with any model, Beamer's direction-optimising rule,
hubs beyond the hub-queue capacity, unreachable vertices and
high-diameter graphs.  The variants are the native ones and the
reference's eight kernel families run on the engine's state.  The frontier features
reduced on the device are checked against the same quantities computed on
the host from the oracle's depths.
"""
import ctypes as C

import numpy as np
import pytest

from belewitte_amd import _native as N
from belewitte_amd import bfs_native, device, graph, switch
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _many_hubs():
    """Root -> 4200 hubs, each with 4100 out-edges: more hubs than the hub
    queue holds (4096), so the rest go through the normal queue."""
    rng = np.random.default_rng(3)
    nh, deg, pool = 4200, 4100, 60000
    V = 1 + nh + pool
    src = [np.zeros(nh, dtype=np.uint32)]
    dst = [np.arange(1, nh + 1, dtype=np.uint32)]
    for h in range(1, nh + 1):
        t = rng.choice(pool, deg, replace=False).astype(np.uint32) + np.uint32(1 + nh)
        src.append(np.full(deg, h, dtype=np.uint32))
        dst.append(t)
    return graph.from_edges(V, np.concatenate(src), np.concatenate(dst))


def _permuted_rmat():
    """An R-MAT with randomly relabelled vertices: the hubs are no longer the
    lowest ids, so a vertex's likely parent (its in-neighbour of largest
    out-degree) sits anywhere in its in-list."""
    g = graph.rmat(12, 16, seed=23)
    perm = np.random.default_rng(5).permutation(g.vertex_count).astype(np.uint32)
    s, d = g.edges()
    p = graph.from_edges(g.vertex_count, perm[s], perm[d])
    p.name = "rmat-12-permuted"
    return p


def _graphs():
    yield graph.rmat(12, 16, seed=21)
    yield _permuted_rmat()
    yield graph.rmat(11, 8, seed=22, undirected=True)
    yield graph.erdos_renyi(5000, 20000, seed=7, undirected=True)
    yield graph.grid2d(64, 48, rewire=0.01, seed=9)
    yield graph.chain(2000)
    yield graph.star(9000)
    yield graph.star(9000, undirected=True)
    yield graph.mesh(20, 30)


@pytest.fixture(scope="module")
def cases(gpu):
    out = []
    for g in _graphs():
        csr = O.CSR.from_graph(g)
        dg = device.DeviceGraph(g)
        eng = bfs_native.BFSEngine(dg, edge_lists=True)
        roots = sorted({g.max_out_degree_vertex(), 0, g.vertex_count // 2, g.vertex_count - 1})
        out.append((g, csr, dg, eng, roots))
    yield out
    for _, _, dg, eng, _ in out:
        eng.close()
        dg.free()


def _want(csr, root):
    levels, nlev = O.bfs_queue(csr, root)
    _, counts = O.bfs_run(csr, root, [4])
    assert counts.size == nlev
    return levels, counts


@pytest.mark.parametrize("variant", bfs_native.VARIANTS)
@pytest.mark.parametrize("use_graph", [False, True])
def test_fixed_variant_bit_exact(cases, variant, use_graph):
    for g, csr, dg, eng, roots in cases:
        for root in roots:
            print(f"case {g.name} root={root} variant={variant} graph={use_graph}", flush=True)
            r = eng.run(root, "fixed", variant, use_graph=use_graph, levels_per_sync=4 if use_graph else 1)
            want, counts = _want(csr, root)
            assert np.array_equal(r.levels, want), (g.name, root, variant)
            assert np.array_equal(r.counts, counts), (g.name, root, variant)
            assert np.all(r.choices == bfs_native.variant_id(variant))


@pytest.mark.parametrize("schedule", [["td", "bu"], ["bu-wave", "td", "bu"], ["bu", "bu", "td", "bu-wave"]])
def test_schedules_bit_exact(cases, schedule):
    for g, csr, dg, eng, roots in cases:
        root = roots[0]
        r = eng.run(root, "schedule", schedule=schedule, levels_per_sync=8)
        want, counts = _want(csr, root)
        assert np.array_equal(r.levels, want), (g.name, schedule)
        assert np.array_equal(r.counts, counts), (g.name, schedule)
        ids = [bfs_native.variant_id(v) for v in schedule]
        assert r.choices.tolist() == [ids[d % len(ids)] for d in range(r.num_levels)]


def test_random_model_never_changes_depths(cases):
    for seed in range(3):
        m = switch.Model.init(np.random.default_rng(seed).normal(8, 3, (64, bfs_native.NUM_FEATURES)), seed=seed, target_mean=10.0)
        for g, csr, dg, eng, roots in cases:
            eng.set_model(m.w)
            for root in roots[:2]:
                r = eng.run(root, "learned", "td", levels_per_sync=8)
                want, counts = _want(csr, root)
                assert np.array_equal(r.levels, want), (g.name, seed)
                assert np.array_equal(r.counts, counts), (g.name, seed)


def test_many_hubs(gpu):
    g = _many_hubs()
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    try:
        want, counts = _want(csr, 0)
        for v in bfs_native.VARIANTS:
            r = eng.run(0, "fixed", v)
            assert np.array_equal(r.levels, want), v
            assert np.array_equal(r.counts, counts), v
    finally:
        eng.close()
        dg.free()


def test_device_features_match_host(cases):
    lib = N.hip_lib()
    for g, csr, dg, eng, roots in cases[:4]:
        root = roots[0]
        levels, _ = O.bfs_queue(csr, root)
        deg = np.diff(g.offsets()).astype(np.float64)
        V = g.vertex_count
        N.check_hip(lib.bw_bfs_engine_begin(eng._e, root, bfs_native.POLICIES["host"], 0, 1, None))
        d = 0
        while True:
            st = eng.status()
            if st["done"]:
                break
            assert st["depth"] == d
            f = st["features"]
            fr = deg[levels == d]
            n = fr.size
            mean = fr.mean() if n else 0.0
            std = np.sqrt(max((fr ** 2).mean() - mean * mean, 0.0)) if n else 0.0
            lv, le = np.log2(V), np.log2(max(g.edge_count, 1))
            explored = deg[(levels <= d) & (levels != bfs_native.UNVISITED)].sum()
            prevn = np.count_nonzero(levels == d - 1) if d else 0
            want = [np.log2(1 + n) - lv, np.log2(1 + fr.sum()) - le,
                    np.log2(1 + V - np.count_nonzero(levels <= d)) - lv, d, np.log2(1 + mean), np.log2(1 + std)]
            assert np.allclose(f[:6], want, rtol=1e-5, atol=1e-5), (g.name, d, f, want)
            assert f[6] == (-1 if d == 0 else (d - 1) % 3)
            assert (f[7] == 0) == (d == 0)
            want = [lv, np.log2(1 + g.edge_count - explored) - le, np.log2(1 + n) - np.log2(1 + prevn)]
            assert np.allclose(f[8:], want, rtol=1e-5, atol=1e-5), (g.name, d, f, want)
            N.check_hip(lib.bw_bfs_engine_force(eng._e, d % 3, None))
            N.check_hip(lib.bw_bfs_engine_level(eng._e, bfs_native.POLICIES["host"], 0, 1, None))
            d += 1


def test_device_choice_matches_host_model(cases):
    """The MLP evaluated inside the level loop picks what the inference
    kernel picks on the same features."""
    lib = N.hip_lib()
    X = np.random.default_rng(1).normal(10, 4, (256, bfs_native.NUM_FEATURES)).astype(np.float32)
    m = switch.Model.init(X, seed=5, target_mean=12.0)
    for g, csr, dg, eng, roots in cases[:4]:
        eng.set_model(m.w)
        N.check_hip(lib.bw_bfs_engine_begin(eng._e, roots[0], bfs_native.POLICIES["learned"], 0, 1, None))
        while True:
            st = eng.status()
            if st["done"]:
                break
            pred = m.predict(st["features"][None, :])[0]
            assert st["choice"] == int(np.argmin(pred)), (g.name, st["depth"], pred)
            N.check_hip(lib.bw_bfs_engine_level(eng._e, bfs_native.POLICIES["learned"], 0, 1, None))


def _beamer_choices(g, levels):
    """Beamer's rule per level (alpha 14, beta 24), from the oracle's depths:
    after top-down, bottom-up when m_f > m_u / 14 (m_u: E minus the
    out-degrees of every vertex reached so far); after bottom-up, top-down
    when the frontier shrank below V / 24."""
    TD, BU = bfs_native.variant_id("td"), bfs_native.variant_id("bu-wave")
    deg = np.diff(g.offsets()).astype(np.int64)
    reached = levels != bfs_native.UNVISITED
    depth = levels[reached]
    nd = int(depth.max()) + 1
    nf = np.bincount(depth, minlength=nd)
    mf = np.bincount(depth, weights=deg[reached], minlength=nd)
    explored = np.cumsum(mf)
    out, prev, prevn = [], -1, 0
    for d in range(nd):
        if prev != BU:
            c = BU if mf[d] > (g.edge_count - explored[d]) / 14.0 else TD
        else:
            c = BU if (nf[d] >= prevn or nf[d] > g.vertex_count / 24.0) else TD
        out.append(c)
        prev, prevn = c, nf[d]
    return out


def test_beamer_policy(cases):
    """The direction-optimising baseline: exact depths and counts, and the
    per-level choices the rule gives on the oracle's frontiers."""
    for g, csr, dg, eng, roots in cases:
        for root in roots:
            print(f"case {g.name} root={root} beamer", flush=True)
            r = eng.run(root, "beamer", "td", levels_per_sync=8)
            want, counts = _want(csr, root)
            assert np.array_equal(r.levels, want), (g.name, root)
            assert np.array_equal(r.counts, counts), (g.name, root)
            assert r.choices.tolist() == _beamer_choices(g, want), (g.name, root)


def test_edge_list_variants_need_their_representation(gpu):
    """Without the edge lists resident the edge-list families are refused
    as fixed choices and never picked by the learned policy."""
    g = graph.rmat(10, 8, seed=31)
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg)
    try:
        assert eng.available == bfs_native.NATIVE_VARIANTS + bfs_native.REFERENCE_VARIANTS[4:]
        with pytest.raises(N.BelewitteError):
            eng.run(0, "fixed", "edge-list")
        # A model that predicts the edge-list family fastest at every level.
        m = switch.Model.init(np.random.default_rng(0).normal(8, 3, (64, bfs_native.NUM_FEATURES)), seed=1, target_mean=10.0)
        L = switch._layout(m.K)
        m.w[L["W3"]:L["b3"]] = 0.0
        m.w[L["b3"]:L["end"]] = 10.0
        m.w[L["b3"] + bfs_native.variant_id("edge-list")] = 1.0
        m.w[L["b3"] + bfs_native.variant_id("vertex-pull")] = 2.0
        eng.set_model(m.w)
        want, counts = _want(csr, 0)
        r = eng.run(0, "learned", "td")
        assert np.array_equal(r.levels, want) and np.array_equal(r.counts, counts)
        assert np.all(r.choices == bfs_native.variant_id("vertex-pull"))
        ev = bfs_native.BFSEngine(dg, edge_lists=True)
        ev.set_model(m.w)
        r = ev.run(0, "learned", "td")
        assert np.array_equal(r.levels, want)
        assert np.all(r.choices == bfs_native.variant_id("edge-list"))
        ev.close()
    finally:
        eng.close()
        dg.free()


def _depth_model(ref_variant: str, from_depth: int) -> np.ndarray:
    """An 11-output model that predicts `td` fastest below `from_depth` and
    the reference family `ref_variant` from there on: hidden unit 0 carries
    the depth feature through both layers, `td`'s output is 5, the
    family's 5 + 1.5 * (from_depth - 0.5 - depth), every other 10."""
    K, F, H = len(bfs_native.VARIANTS), bfs_native.NUM_FEATURES, switch.HIDDEN
    L = switch._layout(K)
    w = np.zeros(L["end"], dtype=np.float32)
    w[L["std"]:L["std"] + F] = 1.0
    w[L["W1"] + 0 * F + 3] = 1.0   # feature 3 is the depth
    w[L["W2"] + 0 * H + 0] = 1.0
    k = bfs_native.variant_id(ref_variant)
    w[L["W3"] + k * H + 0] = -1.5
    w[L["b3"]:L["end"]] = 10.0
    w[L["b3"] + bfs_native.variant_id("td")] = 5.0
    w[L["b3"] + k] = 5.0 + 1.5 * (from_depth - 0.5)
    return w


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("ref_variant,from_depth", [("vertex-push", 2), ("edge-list", 3), ("vertex-pull-warp", 1)])
def test_learned_launches_reference_kernel_when_picked(cases, use_graph, ref_variant, from_depth):
    """The learned policy launches only the native level kernel until a level
    picks a reference family; that level stalls and the search goes on with
    both kinds.  Depths and counts match the oracle, and every level ran the
    variant the model picked for it."""
    eng_model = _depth_model(ref_variant, from_depth)
    for g, csr, dg, eng, roots in cases[:4]:
        eng.set_model(eng_model)
        for root in roots:
            print(f"case {g.name} root={root} {ref_variant} from depth {from_depth} graph={use_graph}", flush=True)
            want, counts = _want(csr, root)
            for _ in range(2):
                r = eng.run(root, "learned", "td", use_graph=use_graph, levels_per_sync=4)
                assert np.array_equal(r.levels, want) and np.array_equal(r.counts, counts), (g.name, root)
                expect = [bfs_native.variant_id("td") if d < from_depth else bfs_native.variant_id(ref_variant)
                          for d in range(r.num_levels)]
                assert r.choices.tolist() == expect, (g.name, root, r.choices.tolist())


def test_native_model_still_loads(cases):
    """A model with one output per native variant picks among those."""
    g, csr, dg, eng, roots = cases[0]
    m = switch.Model.init(np.random.default_rng(2).normal(8, 3, (64, bfs_native.NUM_FEATURES)), outputs=3, seed=3, target_mean=10.0)
    eng.set_model(m.w)
    r = eng.run(roots[0], "learned", "td")
    want, counts = _want(csr, roots[0])
    assert np.array_equal(r.levels, want) and np.array_equal(r.counts, counts)
    assert np.all(r.choices < 3)


def test_exhaustive_records(cases):
    for g, csr, dg, eng, roots in cases[:5]:
        root = roots[-1]
        levels, counts, recs = eng.exhaustive(root, repeats=1)
        want, wcounts = _want(csr, root)
        assert np.array_equal(levels, want)
        assert np.array_equal(counts, wcounts)
        assert len(recs) == counts.size
        assert [r.depth for r in recs] == list(range(len(recs)))
        for r in recs:
            assert r.ms.shape == (len(bfs_native.VARIANTS),) and np.all(r.ms > 0)


def _union(graphs):
    """Disjoint union of `graphs` (vertex ids shifted) and the shifts."""
    src, dst, base = [], [], [0]
    for g in graphs:
        s, d = g.edges()
        src.append(s.astype(np.uint64) + base[-1])
        dst.append(d.astype(np.uint64) + base[-1])
        base.append(base[-1] + g.vertex_count)
    u = graph.from_edges(base[-1], np.concatenate(src).astype(np.uint32), np.concatenate(dst).astype(np.uint32))
    return u, base


@pytest.mark.parametrize("variant", bfs_native.VARIANTS + ("learned",))
def test_multi_root_union_matches_per_graph_bfs(gpu, variant):
    rng = np.random.default_rng(17)
    graphs = [graph.rmat_n(int(2 ** rng.uniform(6, 11)), int(rng.integers(4, 33)) * 64, seed=100 + i)
              for i in range(24)]
    u, base = _union(graphs)
    dg = device.DeviceGraph(u)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    roots = [base[i] + g.max_out_degree_vertex() for i, g in enumerate(graphs)]
    if variant == "learned":
        eng.set_model(switch.Model.init(rng.normal(6, 3, (64, bfs_native.NUM_FEATURES)), seed=2, target_mean=8.0).w)
        r = eng.run(roots, "learned", "td")
    else:
        r = eng.run(roots, "fixed", variant)
    want_counts = np.zeros(r.num_levels, dtype=np.int64)
    for i, g in enumerate(graphs):
        want, _ = O.bfs_queue(O.CSR.from_graph(g), roots[i] - base[i])
        assert np.array_equal(r.levels[base[i]:base[i + 1]], want), (variant, i)
        reached = want[want != bfs_native.UNVISITED]
        assert reached.max() + 1 <= r.num_levels
        want_counts[:reached.max() + 1] += np.bincount(reached, minlength=reached.max() + 1)
    # counts[d] = discovered at level d = vertices at depth d + 1.
    assert np.array_equal(r.counts[:-1], want_counts[1:r.num_levels]), variant
    assert r.counts[-1] == 0
    eng.close()
    dg.free()


@pytest.mark.parametrize("hub,wave", [(256, 1), (512, 8), (100000, 1000)])
def test_tuning_never_changes_depths(cases, hub, wave):
    for g, csr, dg, eng, roots in cases[:3] + cases[5:7]:
        eng.set_tuning(hub, wave)
        try:
            for v in bfs_native.VARIANTS:
                r = eng.run(roots[0], "fixed", v)
                want, counts = _want(csr, roots[0])
                assert np.array_equal(r.levels, want), (g.name, hub, wave, v)
                assert np.array_equal(r.counts, counts), (g.name, hub, wave, v)
        finally:
            eng.set_tuning()


@pytest.mark.parametrize("lds_queue,hub_cap", [(256, 256), (512, 1024), (8192, 16384)])
def test_queue_sizes_never_change_depths(cases, lds_queue, hub_cap):
    """Tiny LDS queues overflow into direct global appends and a tiny hub
    queue sends most hubs through the normal queue; the largest sizes
    shrink the level kernel's occupancy.  Depths and counts stay exact."""
    for g, csr, dg, eng, roots in cases[:3] + cases[5:7]:
        eng.set_queues(lds_queue, hub_cap)
        eng.set_tuning(256, 8)  # many hubs
        try:
            for v in bfs_native.VARIANTS + ("learned",):
                eng.set_model(switch.Model.init(np.random.default_rng(4).normal(8, 3, (64, bfs_native.NUM_FEATURES)),
                                                seed=4, target_mean=10.0).w)
                r = eng.run(roots[0], "learned" if v == "learned" else "fixed", "td" if v == "learned" else v)
                want, counts = _want(csr, roots[0])
                assert np.array_equal(r.levels, want), (g.name, lds_queue, hub_cap, v)
                assert np.array_equal(r.counts, counts), (g.name, lds_queue, hub_cap, v)
        finally:
            eng.set_queues()
            eng.set_tuning()


def test_many_hubs_small_hub_queue(gpu):
    g = _many_hubs()
    csr = O.CSR.from_graph(g)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg)
    try:
        want, counts = _want(csr, 0)
        for q, c in ((256, 256), (8192, 16384)):
            eng.set_queues(q, c)
            for v in ("td", "bu-wave", "vertex-push"):
                r = eng.run(0, "fixed", v)
                assert np.array_equal(r.levels, want), (q, c, v)
                assert np.array_equal(r.counts, counts), (q, c, v)
    finally:
        eng.close()
        dg.free()


def test_level_limit_is_an_error(gpu):
    g = graph.chain(100)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg, max_levels=10)
    with pytest.raises(N.BelewitteError):
        eng.run(0, "fixed", "td")
    eng.close()
    dg.free()


def test_rejects_bad_arguments(cases):
    g, csr, dg, eng, roots = cases[0]
    with pytest.raises(N.BelewitteError):
        eng.run(g.vertex_count, "fixed", "td")
    with pytest.raises(KeyError):
        eng.run(0, "fixed", "nope")
    fresh = bfs_native.BFSEngine(dg)
    with pytest.raises(N.BelewitteError):
        fresh.run(0, "learned", "td")  # no model set
    for q, c in ((128, 4096), (8256, 4096), (2050, 4096), (2048, 0), (2048, 300), (2048, 16640)):
        with pytest.raises(N.BelewitteError):
            fresh.set_queues(q, c)
    assert fresh.queues == (bfs_native.DEFAULT_LDS_QUEUE, bfs_native.DEFAULT_HUB_CAP)
    fresh.close()
    n = C.c_uint32()
    assert N.hip_lib().bw_bfs_engine_run(eng._e, 0, 9, 0, 1, 1, 0, None, None, None, C.byref(n), None,
                                         None) != 0


@pytest.mark.parametrize("nsrc", [1, 17, 64])
@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("td", [0, 1, 4])
def test_multi_source_bit_exact(cases, nsrc, use_graph, td):
    """td 0: every level bottom-up; 1: every level top-down (the frontier's
    out-edges never exceed E); 4: the direction changes between levels.
    Hub degree 16: every vertex above 16 in-edges goes through the
    edge-parallel hub chunks in the bottom-up levels."""
    rng = np.random.default_rng(nsrc)
    for g, csr, dg, eng, roots in cases:
        ms = bfs_native.MultiSourceBFS(dg)
        ms.set_tuning(32, td, 16 if nsrc == 17 else 128)
        try:
            src = rng.integers(0, g.vertex_count, nsrc)
            src[0] = roots[0]
            if nsrc > 2:
                src[1] = src[2]  # a repeated source
            depths, ms_time, bits = ms.run(src, use_graph=use_graph, levels_per_sync=3)
            total = 0
            for i, s in enumerate(src):
                want, _ = O.bfs_queue(csr, int(s))
                assert np.array_equal(depths[i], want), (g.name, i, int(s))
                total += int(np.count_nonzero(want != bfs_native.UNVISITED)) - 1
            assert int(bits.sum()) == total and bits[-1] == 0
            assert ms_time > 0
        finally:
            ms.close()


@pytest.mark.parametrize("wave,td,hub", [(0, 0, 128), (4, 0, 16), (1 << 20, 0, 1 << 20), (4, 2, 64), (0, 1024, 16),
                                         (64, 0, 16)])
def test_multi_source_tuning(cases, wave, td, hub):
    g, csr, dg, eng, roots = cases[0]
    ms = bfs_native.MultiSourceBFS(dg)
    ms.set_tuning(wave, td, hub)
    src = np.arange(0, g.vertex_count, g.vertex_count // 64)[:64]
    depths, _, _ = ms.run(src)
    for i, s in enumerate(src):
        want, _ = O.bfs_queue(csr, int(s))
        assert np.array_equal(depths[i], want), (wave, td, hub, i)
    with pytest.raises(N.BelewitteError):
        ms.set_tuning(wave, 1025, hub)
    with pytest.raises(N.BelewitteError):
        ms.set_tuning(wave, td, 15)
    ms.close()
