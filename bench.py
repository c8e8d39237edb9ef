"""Benchmark of the hot path on one MI355X.

Workload: PageRank on a directed R-MAT graph, scale 22 (4,194,304 vertices),
edge factor 16 (67,108,864 generated edges, self-loops and duplicates
removed), a/b/c/d = 0.57/0.19/0.19/0.05, seed 2; FP32 ranks, damping 0.85,
uniform 1/N start, dangling mass redistributed uniformly, 30 iterations.
A step is one full 30-iteration run of the fused pull iteration
(vertex-pull-nodiv-fused), inputs resident in HBM.  value = GE/s per
iteration = E / (step time / 30).

Prints one JSON line (rank 0).  Extra fields: the HBM roofline of the
iteration kernels measured with HIP events on their stream; the one-off
preprocessing (relabel, plan) and a first run with it; the CPU baseline
(the oracle's OpenMP pull PageRank, on every core this process may use and
on one pinned core, one warm-up and the median of five, timed over the
reference's computation scope), whose ranks also check the GPU's; the
per-iteration rate of each of the reference's 10 variants; and the BFS
figures: the learned per-level switch against the best fixed variant and
the per-level oracle on an undirected R-MAT scale 25 (GTEPS), the batched
searches (64 sources on R-MAT scale 20; 256 small R-MAT graphs), the
learned switch and the fixed variants on the switch's uniform (Erdos-Renyi
2^24) and grid (4096 x 4096) training graphs, and the reference's
vertex-push on R-MAT scale 16 -- each with its CPU baseline, whose depths
check the GPU's.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "BFS GTEPS: learned switch vs best fixed variant; PageRank GE/s/iter vs HBM"
HBM_PEAK_GBS = 8000.0
SCALE, EDGE_FACTOR, SEED, ITERS = 22, 16, 2, 30


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profile(fused):
    """HBM bytes per iteration measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy kernel) for the kernels `fused` runs,
    from the summary under profiles/ that names the same kernels and launch
    shape; None when there is none."""
    name = "pagerank_binned_rmat_s22_traffic.json" if fused.binning else "pagerank_fused_rmat_s22_traffic.json"
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    try:
        rec = json.loads(p.read_text())
    except (OSError, ValueError):
        return None
    if fused.binning:
        # one record per binned shape the autotuner may choose
        for r in rec.get("shapes", []):
            if tuple(r.get("binning", ())) == tuple(fused.binning):
                return r.get("bytes_per_launch")
        return None
    # The pull record holds one launch shape; another shape's traffic differs.
    if tuple(rec.get("tuning", ())) != tuple(fused.tuning):
        return None
    return rec.get("bytes_per_launch")


def launch_shape(fused, shapes) -> dict:
    if fused.binning:
        r, c, l = fused.binning
        info = fused.bin_info()
        shape = {"binned": True, "rows_per_partition": 1 << r, "chunk_tasks": c, "lane_edges": l,
                 "pairs": info["pairs"], "chunks": info["chunks"], "scatter_blocks": info["blocks"]}
    else:
        shape = {"binned": False, "hot_vertices": fused.tuning[0], "rows_grid": fused.tuning[1],
                 "tiles_per_workgroup": fused.tuning[2], "source_sorted_windows": fused.tuning[3],
                 "window_gathers_per_round": fused.tuning[4]}
    shape["autotune_ms_per_10_iterations"] = shapes
    return shape


def host_cpus():
    """(threads the CPU baselines use, the CPUs this process may run on,
    the box's core count): every allowed CPU, which is the CPU share the
    box gives this process."""
    allowed = sorted(os.sched_getaffinity(0))
    return len(allowed), allowed, os.cpu_count()


class OneCpu:
    """This thread pinned to one CPU, the first it may use, for the
    single-threaded baselines (kernel-runner.cpp:258-283 pins to CPU 0)."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.cpu = min(self.saved)
        os.sched_setaffinity(0, {self.cpu})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.saved)


def median_seconds(run, runs: int = 5, warmup: int = 1):
    """Median of `runs` timings after `warmup` untimed runs; `run` returns
    (result, seconds) and the last result is returned with the median."""
    res = None
    for _ in range(warmup):
        res, _ = run()
    secs = []
    for _ in range(runs):
        res, sec = run()
        secs.append(sec)
    return res, float(np.median(secs)), secs


def cpu_baseline(g, ranks_gpu):
    """The oracle's OpenMP vertex-pull-nodiv (uniform dangling) over the
    reference's computation scope (degrees + iterations,
    pagerank/pagerank.cpp:95-127): all 30 iterations on every allowed core,
    and 5 iterations on one pinned core.  Its ranks check the GPU's."""
    from oracle import binding as O
    csr = O.CSR.from_graph(g)
    threads, allowed, ncpu = host_cpus()
    E = g.edge_count
    ranks, sec, secs = median_seconds(lambda: O.pagerank_pull_omp(csr, ITERS, threads, timed=True))
    one_iters = 5
    with OneCpu() as pin:
        _, sec1, secs1 = median_seconds(lambda: O.pagerank_pull_omp(csr, one_iters, 1, timed=True))
    err = float(np.max(np.abs(ranks.astype(np.float64) - ranks_gpu.astype(np.float64))))
    base = {"value": round(E * ITERS / sec / 1e9, 4), "unit": "GE/s/iter", "cores": threads, "kind": "port",
            "sample": f"full workload: {ITERS} iterations of the oracle's OpenMP vertex-pull-nodiv (uniform "
                      f"dangling) over all {E} edges on {threads} threads (the CPUs this process may use, "
                      f"of {ncpu} on the host); median of 5 after 1 warm-up, {sec:.3f} s, timed over the "
                      f"degrees and iterations",
            "runs_s": [round(x, 4) for x in secs],
            "single_thread": {"value": round(E * one_iters / sec1 / 1e9, 4), "unit": "GE/s/iter", "cores": 1,
                              "pinned_cpu": pin.cpu, "iterations": one_iters,
                              "median_s": round(sec1, 4), "runs_s": [round(x, 4) for x in secs1]},
            "host": {"nproc": ncpu, "allowed_cpus": len(allowed)}}
    return base, err, csr


def oracle_depths(csr, roots, threads: int):
    """Queue-BFS depths of the oracle from each root, `threads` searches at
    a time (the calls release the interpreter lock), and their wall time."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import binding as O
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        out = [d for d, _ in ex.map(lambda r: O.bfs_queue(csr, int(r)), roots)]
    return out, time.perf_counter() - t0


def bfs_cpu_leg(g, sources, teps_div: int):
    """The CPU baseline of a BFS workload and the depths that check the
    GPU: the oracle's OpenMP vertex-push (the reference kernel's level loop)
    from every source on every allowed core, timed over the computation
    scope (bfs/bfs.cpp:78-108); then one source on one pinned core, median
    of 3 after the all-core runs warmed the graph.  Returns (baseline,
    {source: depths})."""
    from oracle import binding as O
    csr = O.CSR.from_graph(g)
    threads, allowed, ncpu = host_cpus()
    deg = np.diff(g.offsets()).astype(np.int64)
    depths, edges, secs = {}, 0, 0.0
    for i, s in enumerate(sources):
        lv, _, sec = O.bfs_vertex_push_omp(csr, int(s), threads, timed=True)
        depths[int(s)] = lv
        edges += int(deg[lv != O.UNVISITED].sum()) // teps_div
        secs += sec
        log(f"cpu bfs source {i + 1}/{len(sources)} ({int(s)}) on {threads} threads: {sec:.3f} s")
    s0 = int(sources[0])

    def one():
        res = O.bfs_vertex_push_omp(csr, s0, 1, timed=True)[::2]
        log(f"cpu bfs source {s0} on one pinned core: {res[1]:.3f} s")
        return res
    with OneCpu() as pin:
        _, sec1, secs1 = median_seconds(one, runs=3, warmup=0)
    e0 = int(deg[depths[s0] != O.UNVISITED].sum()) // teps_div
    conv = "reached out-degree sum / 2 (undirected, Graph500)" if teps_div == 2 else "reached out-degree sum"
    base = {"value": round(edges / secs / 1e9, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"the oracle's OpenMP vertex-push from all {len(sources)} sources on {threads} threads (of "
                      f"{ncpu} on the host), {secs:.2f} s over the level loops; TEPS edges: {conv}",
            "single_thread": {"value": round(e0 / sec1 / 1e9, 4), "unit": "GTEPS", "cores": 1,
                              "pinned_cpu": pin.cpu, "source": s0, "median_s": round(sec1, 4),
                              "runs_s": [round(x, 4) for x in secs1]}}
    return base, depths


def bfs_learned_switch(scale: int, sources: int, samples_out=None, cpu: bool = True) -> dict:
    """Undirected R-MAT (scale, edge factor 16, seed 5), `sources` random
    sources of out-degree >= 1: the learned per-level switch against the
    best fixed variant of all eleven (native and the reference's kernel
    families), the per-level oracle over the same set and Beamer's
    direction-optimising rule, simulated on exhaustive per-level timings
    (from the first 4 sources: every variant at every level, twice) and
    executed end to end on the device from every source; every depth
    checked against the CPU baseline's (the oracle)."""
    from belewitte_amd import bfs_native, device, graph, switch
    t0 = time.perf_counter()
    g = graph.rmat(scale, 16, seed=5, undirected=True)
    log(f"bfs graph {g} built in {time.perf_counter() - t0:.1f}s")
    src = graph.random_sources(g, sources, seed=8)
    base, ref = bfs_cpu_leg(g, src, 2) if cpu else (None, None)
    dg = device.DeviceGraph(g)
    t0 = time.perf_counter()
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    device.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3
    if switch.MODEL_PATH.exists():
        model, origin = switch.Model.load(switch.MODEL_PATH), f"{switch.MODEL_PATH.name} (trained offline on the GPU)"
    else:
        samples, _ = switch.collect(eng, src[:4])
        model, _ = switch.fit(samples)
        origin = "trained in this run on 4 of the sources (no stored model)"
    t0 = time.perf_counter()
    res = switch.evaluate(eng, g, src, model, repeats=2, runs=3, samples_out=samples_out, reference=ref,
                          exhaustive_sources=4, progress=lambda m: log(f"bfs switch: {m}"))
    log(f"bfs switch evaluated in {time.perf_counter() - t0:.1f}s: {res['executed_ms']}")
    # Algorithmic bytes of a top-down pass per source: 8 E_r + 16 V_r.
    deg = np.diff(g.offsets()).astype(np.int64)
    V_r = E_r = 0
    for s in src:
        r = eng.run(int(s), "learned", 0)
        reached = r.levels != bfs_native.UNVISITED
        V_r += int(reached.sum())
        E_r += int(deg[reached].sum())
    t = res["executed_ms"]["learned"] / 1e3
    res["roofline"] = {"bound": "hbm", "algorithmic_bytes": 8 * E_r + 16 * V_r,
                       "achieved_gbs": round((8 * E_r + 16 * V_r) / t / 1e9, 1) if t else None,
                       "peak_gbs": HBM_PEAK_GBS,
                       "note": "top-down byte count; direction-optimising levels touch fewer bytes"}
    res["graph"] = {"name": g.name, "vertices": g.vertex_count, "stored_edges": g.edge_count,
                    "teps_convention": "undirected: reached out-degree sum / 2 (Graph500)"}
    res["engine_setup_ms"] = round(setup_ms, 2)
    res["engine_setup_note"] = ("once per graph, outside the GTEPS figures: the four edge lists the reference's "
                                "edge-list families read (built on the device from the CSR), engine buffers and "
                                "the likely-parent array of the bottom-up probes")
    res["model"] = origin
    res["depth_check"] = "oracle (CPU baseline's depths)" if ref is not None else "none (--no-cpu)"
    if base is not None:
        res["cpu_baseline"] = base
    eng.close()
    dg.free()
    return res


def bfs_training_graphs(sources: int = 16, cpu: bool = True) -> dict:
    """BFS on the two contrasting graphs of the learned switch's training
    set: an undirected Erdos-Renyi graph (2^24 vertices, average degree 16,
    seed 3) and a 4096 x 4096 4-neighbour grid with 1% of its edges rewired
    (seed 4).  For each, executed runs from `sources` random sources (seed
    10) of the learned switch, Beamer's rule and every fixed variant of the
    eleven (device time summed over the sources, GTEPS in the undirected
    convention); the learned runs' depths are checked against the oracle's
    from every source (its parallel queue BFS, timed beside them)."""
    from belewitte_amd import bfs_native, device, graph, switch
    model = switch.Model.load(switch.MODEL_PATH)
    threads, _, _ = host_cpus()
    out = {}
    for g in (graph.erdos_renyi(1 << 24, 134217728, seed=3, undirected=True),
              graph.grid2d(4096, 4096, rewire=0.01, seed=4)):
        t0 = time.perf_counter()
        src = [int(x) for x in graph.random_sources(g, sources, seed=10)]
        ref, ref_s = (None, None)
        if cpu:
            from oracle import binding as O
            ref, ref_s = oracle_depths(O.CSR.from_graph(g), src, threads)
        dg = device.DeviceGraph(g)
        eng = bfs_native.BFSEngine(dg, edge_lists=True)
        eng.set_model(model.w)
        edges, levels, mism = 0, 0, 0
        for i, s in enumerate(src):
            r = eng.run(s, "learned", "td")
            edges += bfs_native.teps_edges(g, r.levels)
            levels += r.num_levels
            if ref is not None and not np.array_equal(r.levels, ref[i]):
                mism += 1
        ms = {}
        for pol, var in [("learned", "td"), ("beamer", "td")] + [("fixed", v) for v in bfs_native.VARIANTS]:
            eng.run(src[0], pol, var, download=False)
            ms[pol if pol != "fixed" else var] = sum(eng.run(s, pol, var, download=False).total_ms for s in src)
        fixed = {k: v for k, v in ms.items() if k not in ("learned", "beamer")}
        best = min(fixed, key=fixed.get)
        out[g.name] = {"vertices": g.vertex_count, "stored_edges": g.edge_count, "sources": len(src),
                       "levels": levels, "teps_edges": edges, "executed_ms": {k: round(v, 3) for k, v in ms.items()},
                       "gteps": {"learned": round(edges / (ms["learned"] / 1e3) / 1e9, 2),
                                 "beamer": round(edges / (ms["beamer"] / 1e3) / 1e9, 2),
                                 "best_fixed": round(edges / (fixed[best] / 1e3) / 1e9, 2)},
                       "best_fixed_variant": best,
                       "speedup_vs_best_fixed": round(fixed[best] / ms["learned"], 3),
                       "speedup_vs_beamer": round(ms["beamer"] / ms["learned"], 3),
                       "depth_mismatches_vs_oracle": mism if ref is not None else None,
                       "oracle_queue_bfs_s": round(ref_s, 3) if ref_s is not None else None}
        log(f"bfs {g.name}: {out[g.name]['executed_ms']} ({time.perf_counter() - t0:.1f}s)")
        eng.close()
        dg.free()
    return out


def bfs_reference_vertex_push(cpu: bool) -> dict:
    """The reference's own BFS workload: vertex-push (all four frontier
    reductions) on a directed R-MAT scale 16, edge factor 16 (seed 1), CSR
    with uint64 offsets, from the max-out-degree vertex, with the reference
    level loop (one frontier read-back per level).  The GPU figure is the
    device time of the level kernels (HIP events per level, summed); the
    Python wall time of the whole call, allocations included, is beside
    it.  Depths are checked against the oracle; the CPU leg is the
    oracle's OpenMP vertex-push over the computation scope, all allowed
    cores and one pinned core, median of 5 after a warm-up."""
    from belewitte_amd import bfs, device, graph
    from oracle import binding as O
    g = graph.rmat(16, 16, seed=1)
    root = g.max_out_degree_vertex()
    dg = device.DeviceGraph(g, offset_bytes=8, reps=("csr",))
    csr = O.CSR.from_graph(g)
    want, _ = O.bfs_queue(csr, root)
    reached = want != 2147483647
    edges = int(np.diff(g.offsets()).astype(np.int64)[reached].sum())
    out = {"graph": {"vertices": g.vertex_count, "edges": g.edge_count, "offset_bytes": 8}, "root": root,
           "teps_convention": "reached out-degree sum (directed)"}
    for name in ("vertex-push", "vertex-push-bulk", "vertex-push-warpreduce", "vertex-push-blockreduce"):
        bfs.run(dg, root, name)
        runs = []
        for _ in range(5):
            t0 = time.perf_counter()
            r = bfs.run(dg, root, name)
            runs.append((float(r.level_ms.sum()), time.perf_counter() - t0))
        dev = float(np.median([x for x, _ in runs]))
        wall = float(np.median([y for _, y in runs]))
        out[name] = {"level_kernels_ms": round(dev, 4), "gteps": round(edges / (dev / 1e3) / 1e9, 3),
                     "call_wall_ms": round(wall * 1e3, 3), "levels": int(r.num_levels),
                     "bit_exact": bool(np.array_equal(r.levels, want))}
    # The same search by the device-driven engine (uint32 offsets, learned
    # switch, level loop on the device).
    from belewitte_amd import bfs_native, switch
    dg32 = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg32)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    eng.run(root, "learned", "td")
    r = min((eng.run(root, "learned", "td") for _ in range(5)), key=lambda x: x.total_ms)
    out["native_engine_learned"] = {"ms": round(r.total_ms, 4), "gteps": round(edges / (r.total_ms / 1e3) / 1e9, 3),
                                    "bit_exact": bool(np.array_equal(r.levels, want)), "levels": r.num_levels}
    eng.close()
    dg32.free()
    if cpu:
        threads, _, ncpu = host_cpus()
        (lv, _), sec, secs = median_seconds(
            lambda: (lambda r: ((r[0], r[1]), r[2]))(O.bfs_vertex_push_omp(csr, root, threads, timed=True)))
        out["cpu_baseline"] = {"value": round(edges / sec / 1e9, 4), "unit": "GTEPS", "cores": threads,
                               "kind": "port", "median_ms": round(sec * 1e3, 4),
                               "sample": f"the oracle's OpenMP vertex-push on {threads} threads (of {ncpu}), "
                                         f"median of 5 after 1 warm-up over the level loop",
                               "bit_exact": bool(np.array_equal(lv, want))}
        with OneCpu() as pin:
            (lv1, _), sec1, _ = median_seconds(
                lambda: (lambda r: ((r[0], r[1]), r[2]))(O.bfs_vertex_push_omp(csr, root, 1, timed=True)))
        out["cpu_baseline"]["single_thread"] = {"value": round(edges / sec1 / 1e9, 4), "unit": "GTEPS",
                                                "cores": 1, "pinned_cpu": pin.cpu, "median_ms": round(sec1 * 1e3, 4),
                                                "bit_exact": bool(np.array_equal(lv1, want))}
    dg.free()
    return out


def bfs_batched(cpu: bool = True) -> dict:
    """Batch 1: 64 sources on a directed R-MAT scale 20, edge factor 16
    (seed 6) in one multi-source search; batch 2: 256 directed R-MAT graphs
    of 2^10..2^16 vertices (log-uniform) and edge factors 4..32 searched as
    one disjoint union from their max-out-degree vertices with the learned
    per-level switch (the best fixed variant beside it).  Both autotuned,
    both one device-driven level loop in hipGraph batches; every search's
    depths checked against the oracle's queue BFS (timed beside them)."""
    from belewitte_amd import batch, bfs_native, device, graph, switch
    threads, _, _ = host_cpus()
    out = {}
    g = graph.rmat(20, 16, seed=6)
    dg = device.DeviceGraph(g)
    src = graph.random_sources(g, 64, seed=9)
    ms = bfs_native.MultiSourceBFS(dg)
    tuned = batch.autotune_msbfs(ms, src)
    depths, t, _ = ms.run(src)
    best = min(t, *(ms.run(src, download=False)[1] for _ in range(3)))
    deg = np.diff(g.offsets()).astype(np.int64)
    edges = sum(batch.reached_edges(deg, depths[i]) for i in range(len(src)))
    rec = {"ms": round(best, 3), "gteps": round(edges / (best / 1e3) / 1e9, 2),
           "wave_degree": tuned.params[0], "topdown_divisor": tuned.params[1], "hub_degree": tuned.params[2],
           "tuning_ms": tuned.table, "reached_edges": edges}
    if cpu:
        from oracle import binding as O
        ref, sec = oracle_depths(O.CSR.from_graph(g), src, threads)
        rec["depth_mismatches_vs_oracle"] = sum(int(not np.array_equal(depths[i], ref[i])) for i in range(len(src)))
        rec["cpu_oracle_queue_bfs"] = {"s": round(sec, 3), "threads": threads,
                                       "gteps": round(edges / sec / 1e9, 3)}
    out["sources_64_rmat_s20"] = rec
    ms.close()
    dg.free()

    graphs = batch.small_rmat_batch(256, seed=7)
    u, base = batch.union(graphs)
    roots = [base[i] + gi.max_out_degree_vertex() for i, gi in enumerate(graphs)]
    dg = device.DeviceGraph(u)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    # The best fixed variant, for contrast with the learned per-level switch.
    fixed = {v: min(eng.run(roots, "fixed", v, download=False).total_ms for _ in range(3))
             for v in bfs_native.VARIANTS}
    tuned = batch.autotune_engine(eng, roots, "learned", "td")
    best = min(eng.run(roots, "learned", "td", download=False).total_ms for _ in range(5))
    r = eng.run(roots, "learned", "td")
    deg = np.diff(u.offsets()).astype(np.int64)
    edges = batch.reached_edges(deg, r.levels)
    rec = {"ms": round(best, 3), "gteps": round(edges / (best / 1e3) / 1e9, 2),
           "policy": "learned per-level switch (shipped model)", "choices": r.choices.tolist(),
           "fixed_variant_ms": {k: round(v, 3) for k, v in fixed.items()},
           "vertices": u.vertex_count, "edges": u.edge_count, "reached_edges": edges,
           "levels": r.num_levels,
           "tuned": dict(zip(("hub_degree", "wave_degree", "lds_queue", "hub_cap"), tuned.params)),
           "tuning_ms": tuned.table}
    if cpu:
        from concurrent.futures import ThreadPoolExecutor

        from oracle import binding as O
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            ref = list(ex.map(lambda gi: O.bfs_queue(O.CSR.from_graph(gi), gi.max_out_degree_vertex())[0], graphs))
        sec = time.perf_counter() - t0
        rec["depth_mismatches_vs_oracle"] = sum(
            int(not np.array_equal(r.levels[base[i]:base[i] + gi.vertex_count], ref[i])) for i, gi in enumerate(graphs))
        rec["cpu_oracle_queue_bfs"] = {"s": round(sec, 3), "threads": threads}
    out["graphs_256_rmat"] = rec
    eng.close()
    dg.free()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-variants", action="store_true", help="skip the per-variant table")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-bfs", action="store_true", help="skip the BFS figures")
    ap.add_argument("--no-training-graphs", action="store_true",
                    help="skip the BFS runs on the Erdos-Renyi and grid graphs")
    ap.add_argument("--bfs-scale", type=int, default=25, help="R-MAT scale of the learned-switch BFS")
    ap.add_argument("--bfs-sources", type=int, default=16)
    ap.add_argument("--bfs-samples-out", default=None, help="save the exhaustive per-level timings (.npz)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the graph kernels "
              "here run on a single GPU and no multi-GPU path is built.", file=sys.stderr)
        return 2

    from belewitte_amd import device, graph, pagerank
    device.require_gpu()
    device.set_device(0)
    props = device.device_props(0)

    t0 = time.perf_counter()
    g = graph.rmat(SCALE, EDGE_FACTOR, seed=SEED)
    V, E = g.vertex_count, g.edge_count
    log(f"graph {g} built in {time.perf_counter() - t0:.1f}s")
    # One-off preprocessing, each step timed: the upload of the graph as
    # generated, its relabelling by out-degree on the device, the plan (pull
    # tiles and degrees, then the default binned shape, built on the device),
    # and a first 30-iteration run on top: the end-to-end time of a single
    # cold run from the generated order.
    def timed(fn):
        device.synchronize()
        t = time.perf_counter()
        r = fn()
        device.synchronize()
        return r, time.perf_counter() - t

    src, t_upload = timed(lambda: device.DeviceGraph(g, reps=("csr", "rev_csr")))
    (rdg, perm), t_relabel = timed(src.relabel_by_degree)
    fused, t_create = timed(lambda: pagerank.FusedPageRank(rdg, max_iterations=ITERS))
    fused.perm, fused._owns_dg = perm, True
    _, t_bin = timed(lambda: fused.bin(14, 128, 8))
    _, t_run = timed(lambda: fused.run(ITERS, "uniform", download=False))
    t_host_relabel = timed(g.reorder_by_degree)[1]
    prep_ms = {"upload_generated_order_ms": round(t_upload * 1e3, 2),
               "relabel_on_device_ms": round(t_relabel * 1e3, 2),
               "plan_create_ms": round(t_create * 1e3, 2), "binned_plan_ms": round(t_bin * 1e3, 2),
               "first_run_30_iterations_ms": round(t_run * 1e3, 2)}
    t_e2e = t_relabel + t_create + t_bin + t_run
    # The same steps again on the same graph: the cost per further graph,
    # once the process has loaded the kernels (the first calls above include
    # the one-time loading of the sort and scan kernels).
    (rdg2, _), w_relabel = timed(src.relabel_by_degree)
    fused2, w_create = timed(lambda: pagerank.FusedPageRank(rdg2, max_iterations=ITERS))
    fused2._owns_dg = True
    _, w_bin = timed(lambda: fused2.bin(14, 128, 8))
    _, w_run = timed(lambda: fused2.run(ITERS, "uniform", download=False))
    fused2.close()
    prep_ms["warm"] = {"relabel_on_device_ms": round(w_relabel * 1e3, 2), "plan_create_ms": round(w_create * 1e3, 2),
                       "binned_plan_ms": round(w_bin * 1e3, 2), "run_30_iterations_ms": round(w_run * 1e3, 2)}
    log(f"preprocessing on the device: {prep_ms}; host relabel for comparison {t_host_relabel:.2f}s")
    shapes = fused.autotune()
    log(f"launch shapes (ms per 10 iterations): {shapes}; using {fused.kernel_names()}")

    # Warm-up, then K timed steps bracketed by device synchronisation.
    for _ in range(args.warmup):
        fused.run(ITERS, "uniform", download=False)
    device.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, _, ms = fused.run(ITERS, "uniform", download=False)
        kernel_ms.append(ms)
    device.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    iter_s = ms_per_step / 1e3 / ITERS
    value = E / iter_s / 1e9

    # Roofline of one iteration (its two launches): algorithmic bytes
    # 8E + 16V + 4 of the pull iteration over the average iteration
    # duration, HIP events on the launches' stream.
    bytes_per_launch = 8 * E + 16 * V + 4
    avg_launch_s = float(np.mean(kernel_ms)) / 1e3 / ITERS
    achieved = bytes_per_launch / avg_launch_s / 1e9
    log(f"step {ms_per_step:.3f} ms, launch avg {avg_launch_s * 1e6:.1f} us, {achieved:.0f} GB/s")

    ranks, resid, _ = fused.run(ITERS, "uniform")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GE/s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pagerank vertex-pull-nodiv-fused, directed R-MAT scale 22 edge factor 16 "
                               "(a,b,c=0.57,0.19,0.19, seed 2), 30 iterations, uniform dangling, "
                               "vertices relabelled by out-degree on the device",
                   "vertices": V, "edges": E, "iterations": ITERS, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(fused),
                     "kernel": fused.kernel_names() + " (one iteration)",
                     "launch_shape": launch_shape(fused, shapes),
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": round(avg_launch_s * 1e6, 2)},
        "device": {"arch": props["arch"], "compute_units": props["compute_units"]},
        "final_residual": float(resid[-1]),
        "preprocessing": dict(prep_ms, **{
            "plan_total_ms": round((t_create + t_bin) * 1e3, 2),
            "end_to_end_from_generated_order_ms": round(t_e2e * 1e3, 2),
            "end_to_end_ge_s_iter": round(E * ITERS / t_e2e / 1e9, 3),
            "host_relabel_for_comparison_s": round(t_host_relabel, 3),
            "note": "once per graph, outside value: relabelling by out-degree, pull plan and binned plan, all "
                    "on the device from the uploaded graph as generated; end_to_end_* is relabel + plans + a "
                    "cold 30-iteration run (upload excluded); warm: the same steps on a second graph in the "
                    "same process"}),
    }

    if not args.no_cpu:
        base, err_cpu, csr = cpu_baseline(g, ranks)
        out["cpu_baseline"] = base
        from oracle import binding as O
        x = O.pagerank_f64(csr, ITERS, "uniform")
        out["accuracy"] = {"max_abs_vs_cpu_f32": err_cpu,
                           "max_abs_vs_f64": float(np.max(np.abs(ranks - x))),
                           "max_rel_vs_f64": float(np.max(np.abs(ranks - x) / np.maximum(x, 1e-30)))}
        log(f"cpu baseline {base['value']} GE/s/iter, max-abs vs cpu {err_cpu:.3e}")

    # The same fused kernel on the graph in its generated vertex order.
    dg = src
    # Its own autotuned shape: the pairs of the binned layout follow the
    # vertex order, so the relabelled graph's best shape need not be its.
    natural = pagerank.FusedPageRank(dg, max_iterations=ITERS)
    nat_shapes = natural.autotune()
    natural.run(ITERS, "uniform", download=False)
    ms_nat = min(natural.run(ITERS, "uniform", download=False)[2] for _ in range(3))
    nat_s = ms_nat / 1e3 / ITERS
    out["natural_order"] = {"ge_s_iter": round(E / nat_s / 1e9, 3), "avg_launch_us": round(nat_s * 1e6, 2),
                            "frac": round(bytes_per_launch / nat_s / 1e9 / HBM_PEAK_GBS, 4),
                            "shape": natural.kernel_names(), "autotune_ms_per_10_iterations": nat_shapes}
    if natural.binning and fused.binning:
        # The binned layout's size follows the vertex order: a pair is a
        # (source, partition) with an edge between them, and the relabel puts
        # the most-gathered vertices into few partitions.
        out["natural_order"]["pairs"] = natural.bin_info()["pairs"]
        out["natural_order"]["pairs_relabelled"] = fused.bin_info()["pairs"]
    natural.close()

    if not args.no_variants:
        # Each variant through bw_pr_run (degrees, plans and iterations on
        # the device): GE/s per iteration from the difference of a 40- and
        # a 10-iteration run (the one-off setup cancels), best of 3 each,
        # and the wall time of a whole 30-iteration run.
        dg.ensure("csr")
        table, runs = {}, {}
        for name in pagerank.VARIANTS:
            dg.ensure(pagerank.REPRESENTATION[pagerank.variant_id(name)])
            pagerank.run(dg, name, iterations=2, mode="uniform")
            t = {}
            for it in (10, 40, ITERS):
                best = np.inf
                for _ in range(3):
                    device.synchronize()
                    t0 = time.perf_counter()
                    pagerank.run(dg, name, iterations=it, mode="uniform")
                    best = min(best, time.perf_counter() - t0)
                t[it] = best
            per = (t[40] - t[10]) / 30
            table[name] = round(E / per / 1e9, 3)
            runs[name] = round(t[ITERS] * 1e3, 3)
            log(f"variant {name}: {table[name]} GE/s/iter, {runs[name]} ms per 30-iteration run")
        table[pagerank.FUSED] = round(value, 3)
        out["variants_ge_s_iter"] = table
        out["variants_run_ms_30_iterations"] = runs
    fused.close()
    dg.free()

    if not args.no_bfs:
        out["bfs_learned_switch"] = bfs_learned_switch(args.bfs_scale, args.bfs_sources, args.bfs_samples_out,
                                                       cpu=not args.no_cpu)
        out["bfs_batched"] = bfs_batched(cpu=not args.no_cpu)
        if not args.no_training_graphs:
            out["bfs_uniform_and_grid"] = bfs_training_graphs(cpu=not args.no_cpu)
        out["bfs_reference_vertex_push_rmat_s16"] = bfs_reference_vertex_push(not args.no_cpu)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
