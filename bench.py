"""Benchmark of the hot path on one MI355X.

Workload: PageRank on a directed R-MAT graph, scale 22 (4,194,304 vertices),
edge factor 16 (67,108,864 generated edges, self-loops and duplicates
removed), a/b/c/d = 0.57/0.19/0.19/0.05, seed 2; FP32 ranks, damping 0.85,
uniform 1/N start, dangling mass redistributed uniformly, 30 iterations.
A step is one full 30-iteration run of the fused pull iteration
(vertex-pull-nodiv-fused), inputs resident in HBM.  value = GE/s per
iteration = E / (step time / 30).

Prints one JSON line (rank 0).  Extra fields: the HBM roofline of the
iteration kernel measured with HIP events on its stream, the CPU baseline
(the oracle's OpenMP pull PageRank on the host cores, the same workload),
the error of the GPU ranks against that baseline and against the FP64
oracle, the per-iteration rate of each of the reference's 10 variants, and
the BFS figures: the learned per-level switch against the best fixed
variant and the per-level oracle on an undirected R-MAT scale 25 (GTEPS,
with a CPU baseline), the batched searches (64 sources on R-MAT scale 20;
256 small R-MAT graphs), the learned switch and the fixed variants on the
switch's uniform (Erdos-Renyi 2^24) and grid (4096 x 4096) training graphs,
and the reference's vertex-push on R-MAT scale 16 (GPU and CPU).
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


def cpu_baseline(g, ranks_gpu, threads: int):
    from oracle import binding as O
    csr = O.CSR.from_graph(g)
    t0 = time.perf_counter()
    ranks = O.pagerank_pull_omp(csr, ITERS, threads)
    dt = time.perf_counter() - t0
    ge_s = g.edge_count * ITERS / dt / 1e9
    err = float(np.max(np.abs(ranks.astype(np.float64) - ranks_gpu.astype(np.float64))))
    return {"value": round(ge_s, 4), "unit": "GE/s/iter", "cores": threads, "kind": "port",
            "sample": f"full workload: {ITERS} iterations of the oracle's OpenMP vertex-pull-nodiv "
                      f"(uniform dangling) over all {g.edge_count} edges, {dt:.2f} s"}, err, csr


def bfs_learned_switch(scale: int, sources: int, samples_out=None) -> dict:
    """Undirected R-MAT (scale, edge factor 16, seed 5), `sources` random
    sources of out-degree >= 1: the learned per-level switch against the
    best fixed native variant and the per-level oracle, simulated on
    exhaustive per-level timings and executed end to end on the device."""
    from belewitte_amd import bfs_native, device, graph, switch
    t0 = time.perf_counter()
    g = graph.rmat(scale, 16, seed=5, undirected=True)
    log(f"bfs graph {g} built in {time.perf_counter() - t0:.1f}s")
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg)
    src = graph.random_sources(g, sources, seed=8)
    if switch.MODEL_PATH.exists():
        model, origin = switch.Model.load(switch.MODEL_PATH), f"{switch.MODEL_PATH.name} (trained offline on the GPU)"
    else:
        samples, _ = switch.collect(eng, src[:4])
        model, _ = switch.fit(samples)
        origin = "trained in this run on 4 of the sources (no stored model)"
    t0 = time.perf_counter()
    res = switch.evaluate(eng, g, src, model, repeats=2, runs=3, samples_out=samples_out)
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
                       "peak_gbs": HBM_PEAK_GBS}
    res["graph"] = {"name": g.name, "vertices": g.vertex_count, "stored_edges": g.edge_count,
                    "teps_convention": "undirected: reached out-degree sum / 2 (Graph500)"}
    res["model"] = origin
    eng.close()
    dg.free()
    return res


def bfs_training_graphs(sources: int = 16) -> dict:
    """BFS on the two contrasting graphs of the learned switch's training
    set: an undirected Erdos-Renyi graph (2^24 vertices, average degree 16,
    seed 3) and a 4096 x 4096 4-neighbour grid with 1% of its edges rewired
    (seed 4).  For each, executed runs from `sources` random sources (seed
    10) of the learned switch and of every fixed native variant (device
    time summed over the sources, GTEPS in the undirected convention); the
    learned runs' depths are compared with the fixed top-down runs' on two
    sources."""
    from belewitte_amd import bfs_native, device, graph, switch
    model = switch.Model.load(switch.MODEL_PATH)
    out = {}
    for g in (graph.erdos_renyi(1 << 24, 134217728, seed=3, undirected=True),
              graph.grid2d(4096, 4096, rewire=0.01, seed=4)):
        t0 = time.perf_counter()
        dg = device.DeviceGraph(g)
        eng = bfs_native.BFSEngine(dg)
        eng.set_model(model.w)
        src = [int(x) for x in graph.random_sources(g, sources, seed=10)]
        edges, levels, same = 0, 0, True
        for i, s in enumerate(src):
            r = eng.run(s, "learned", "td")
            edges += bfs_native.teps_edges(g, r.levels)
            levels += r.num_levels
            if i < 2:
                same &= bool(np.array_equal(r.levels, eng.run(s, "fixed", "td").levels))
        ms = {}
        for pol, var in [("learned", "td")] + [("fixed", v) for v in bfs_native.VARIANTS]:
            eng.run(src[0], pol, var, download=False)
            ms["learned" if pol == "learned" else var] = sum(
                eng.run(s, pol, var, download=False).total_ms for s in src)
        fixed = {k: v for k, v in ms.items() if k != "learned"}
        best = min(fixed, key=fixed.get)
        out[g.name] = {"vertices": g.vertex_count, "stored_edges": g.edge_count, "sources": len(src),
                       "levels": levels, "teps_edges": edges, "executed_ms": {k: round(v, 3) for k, v in ms.items()},
                       "gteps": {"learned": round(edges / (ms["learned"] / 1e3) / 1e9, 2),
                                 "best_fixed": round(edges / (fixed[best] / 1e3) / 1e9, 2)},
                       "best_fixed_variant": best,
                       "speedup_vs_best_fixed": round(fixed[best] / ms["learned"], 3),
                       "depths_equal_to_fixed_td": same}
        log(f"bfs {g.name}: {out[g.name]['executed_ms']} ({time.perf_counter() - t0:.1f}s)")
        eng.close()
        dg.free()
    return out


def bfs_reference_vertex_push(cpu: bool) -> dict:
    """The reference's own BFS workload: vertex-push (all four frontier
    reductions) on a directed R-MAT scale 16, edge factor 16 (seed 1), CSR
    with uint64 offsets, from the max-out-degree vertex, with the reference
    level loop (one frontier read-back per level).  Depths are checked
    against the oracle; the CPU leg is the oracle's OpenMP vertex-push."""
    from belewitte_amd import bfs, device, graph
    from oracle import binding as O
    g = graph.rmat(16, 16, seed=1)
    root = g.max_out_degree_vertex()
    dg = device.DeviceGraph(g, offset_bytes=8, reps=("csr",))
    csr = O.CSR.from_graph(g)
    want, _ = O.bfs_queue(csr, root)
    reached = want != 2147483647
    edges = int(np.diff(g.offsets()).astype(np.int64)[reached].sum())
    out = {"graph": {"vertices": g.vertex_count, "edges": g.edge_count, "offset_bytes": 8}, "root": root}
    for name in ("vertex-push", "vertex-push-bulk", "vertex-push-warpreduce", "vertex-push-blockreduce"):
        bfs.run(dg, root, name)
        best = np.inf
        for _ in range(5):
            t0 = time.perf_counter()
            r = bfs.run(dg, root, name, time_levels=False)
            best = min(best, time.perf_counter() - t0)
        out[name] = {"ms": round(best * 1e3, 3), "gteps": round(edges / best / 1e9, 3),
                     "bit_exact": bool(np.array_equal(r.levels, want))}
    # The same search by the device-driven engine (uint32 offsets, learned
    # switch, level loop on the device), for contrast with the reference
    # loop's per-level host round trip and thread-per-vertex expansion.
    from belewitte_amd import bfs_native, switch
    dg32 = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg32)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    eng.run(root, "learned", "td")
    r = min((eng.run(root, "learned", "td") for _ in range(5)), key=lambda x: x.total_ms)
    out["native_engine_learned"] = {"ms": round(r.total_ms, 3), "gteps": round(edges / (r.total_ms / 1e3) / 1e9, 3),
                                    "bit_exact": bool(np.array_equal(r.levels, want)), "levels": r.num_levels}
    eng.close()
    dg32.free()
    if cpu:
        threads = min(len(os.sched_getaffinity(0)), 16)
        for th in (threads, 1):
            O.bfs_vertex_push_omp(csr, root, th)
            best = np.inf
            for _ in range(5):
                t0 = time.perf_counter()
                lv, _ = O.bfs_vertex_push_omp(csr, root, th)
                best = min(best, time.perf_counter() - t0)
            out[f"cpu_vertex_push_{th}_threads"] = {"ms": round(best * 1e3, 3),
                                                    "gteps": round(edges / best / 1e9, 3),
                                                    "bit_exact": bool(np.array_equal(lv, want))}
    dg.free()
    return out


def bfs_cpu_baseline(scale: int) -> dict:
    """The oracle's OpenMP vertex-push BFS (the reference kernel's level
    loop restated) on the learned-switch graph from one source: a bounded
    sample of the BFS workload on the host cores."""
    from belewitte_amd import graph
    from oracle import binding as O
    g = graph.rmat(scale, 16, seed=5, undirected=True)
    src = graph.random_sources(g, 1, seed=8)[0]
    csr = O.CSR.from_graph(g)
    threads = min(len(os.sched_getaffinity(0)), 16)
    t0 = time.perf_counter()
    lv, counts = O.bfs_vertex_push_omp(csr, src, threads)
    dt = time.perf_counter() - t0
    edges = int(np.diff(g.offsets()).astype(np.int64)[lv != 2147483647].sum()) // 2
    return {"value": round(edges / dt / 1e9, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": f"one source ({src}) of the 16, {len(counts)} levels, {dt:.2f} s, "
                      f"undirected TEPS convention (reached out-degree sum / 2)"}


def bfs_batched() -> dict:
    """Batch 1: 64 sources on a directed R-MAT scale 20, edge factor 16
    (seed 6) in one multi-source search; batch 2: 256 directed R-MAT graphs
    of 2^10..2^16 vertices (log-uniform) and edge factors 4..32 searched as
    one disjoint union from their max-out-degree vertices with the learned
    per-level switch (the best fixed variant beside it).  Both autotuned,
    both one device-driven level loop in hipGraph batches."""
    from belewitte_amd import batch, bfs_native, device, graph, switch
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
    # Spot check against single-source searches of the engine.
    eng = bfs_native.BFSEngine(dg)
    agree = all(np.array_equal(eng.run(int(src[i]), "fixed", "bu-wave").levels, depths[i]) for i in (0, 31, 63))
    eng.close()
    out["sources_64_rmat_s20"] = {"ms": round(best, 3), "gteps": round(edges / (best / 1e3) / 1e9, 2),
                                  "wave_degree": tuned.params[0], "topdown_divisor": tuned.params[1],
                                  "hub_degree": tuned.params[2],
                                  "tuning_ms": tuned.table,
                                  "reached_edges": edges, "spot_check_vs_single_source": bool(agree)}
    ms.close()
    dg.free()

    graphs = batch.small_rmat_batch(256, seed=7)
    u, base = batch.union(graphs)
    roots = [base[i] + gi.max_out_degree_vertex() for i, gi in enumerate(graphs)]
    dg = device.DeviceGraph(u)
    eng = bfs_native.BFSEngine(dg)
    eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    # The best fixed variant, for contrast with the learned per-level switch.
    fixed = {v: min(eng.run(roots, "fixed", v, download=False).total_ms for _ in range(3))
             for v in bfs_native.VARIANTS}
    tuned = batch.autotune_engine(eng, roots, "learned", "td")
    best = min(eng.run(roots, "learned", "td", download=False).total_ms for _ in range(5))
    r = eng.run(roots, "learned", "td")
    deg = np.diff(u.offsets()).astype(np.int64)
    edges = batch.reached_edges(deg, r.levels)
    out["graphs_256_rmat"] = {"ms": round(best, 3), "gteps": round(edges / (best / 1e3) / 1e9, 2),
                              "policy": "learned per-level switch (shipped model)",
                              "choices": r.choices.tolist(),
                              "fixed_variant_ms": {k: round(v, 3) for k, v in fixed.items()},
                              "vertices": u.vertex_count, "edges": u.edge_count, "reached_edges": edges,
                              "levels": r.num_levels, "tuned_hub_wave": list(tuned.params),
                              "tuning_ms": tuned.table}
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
    t0 = time.perf_counter()
    fused = pagerank.FusedPageRank.for_graph(g, max_iterations=ITERS, reorder=True)
    log(f"degree-ordered upload + plan in {time.perf_counter() - t0:.1f}s: {fused.info()}")
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
                               "vertices relabelled by out-degree before upload",
                   "vertices": V, "edges": E, "iterations": ITERS, "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(fused),
                     "kernel": fused.kernel_names() + " (one iteration)",
                     "launch_shape": launch_shape(fused, shapes),
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": round(avg_launch_s * 1e6, 2)},
        "device": {"arch": props["arch"], "compute_units": props["compute_units"]},
        "final_residual": float(resid[-1]),
    }

    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        threads = min(threads, 16)
        base, err_cpu, csr = cpu_baseline(g, ranks, threads)
        out["cpu_baseline"] = base
        from oracle import binding as O
        x = O.pagerank_f64(csr, ITERS, "uniform")
        out["accuracy"] = {"max_abs_vs_cpu_f32": err_cpu,
                           "max_abs_vs_f64": float(np.max(np.abs(ranks - x))),
                           "max_rel_vs_f64": float(np.max(np.abs(ranks - x) / np.maximum(x, 1e-30)))}
        log(f"cpu baseline {base['value']} GE/s/iter, max-abs vs cpu {err_cpu:.3e}")

    # The same fused kernel on the graph in its generated vertex order.
    dg = device.DeviceGraph(g, reps=("rev_csr",))
    natural = pagerank.FusedPageRank(dg, max_iterations=ITERS)
    natural._apply(("bins",) + fused.binning if fused.binning else fused.tuning)
    natural.run(ITERS, "uniform", download=False)
    _, _, ms_nat = natural.run(ITERS, "uniform", download=False)
    out["natural_order"] = {"ge_s_iter": round(E / (ms_nat / 1e3 / ITERS) / 1e9, 3),
                            "avg_launch_us": round(ms_nat * 1e3 / ITERS, 2)}
    natural.close()

    if not args.no_variants:
        dg.ensure("csr")
        table = {}
        for name in pagerank.VARIANTS:
            dg.ensure(pagerank.REPRESENTATION[pagerank.variant_id(name)])
            device.synchronize()
            t0 = time.perf_counter()
            pagerank.run(dg, name, iterations=ITERS, mode="uniform")
            dt = time.perf_counter() - t0
            table[name] = round(E * ITERS / dt / 1e9, 3)
            log(f"variant {name}: {table[name]} GE/s/iter")
        table[pagerank.FUSED] = round(value, 3)
        out["variants_ge_s_iter"] = table
    fused.close()
    dg.free()

    if not args.no_bfs:
        out["bfs_learned_switch"] = bfs_learned_switch(args.bfs_scale, args.bfs_sources, args.bfs_samples_out)
        if not args.no_cpu:
            out["bfs_learned_switch"]["cpu_baseline"] = bfs_cpu_baseline(args.bfs_scale)
            log(f"bfs cpu baseline {out['bfs_learned_switch']['cpu_baseline']}")
        out["bfs_batched"] = bfs_batched()
        if not args.no_training_graphs:
            out["bfs_uniform_and_grid"] = bfs_training_graphs()
        out["bfs_reference_vertex_push_rmat_s16"] = bfs_reference_vertex_push(not args.no_cpu)

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
