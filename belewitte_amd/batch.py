"""Batched BFS and the autotuner of its work-division thresholds.

Two batch shapes, each run as one device-driven level loop launched in
captured hipGraph batches (no host round trip per level):

* many sources on one graph: :class:`~belewitte_amd.bfs_native.MultiSourceBFS`
  (64 searches share every pass over the graph);
* many small graphs: their disjoint union (:func:`union`) searched once from
  all their roots (:class:`~belewitte_amd.bfs_native.BFSEngine` multi-root);
  each component gets exactly its own BFS depths.

The autotuner times whole batch runs over a grid of the kernels'
thread/wave/workgroup thresholds and queue sizes and keeps the fastest:
- the engine: hub out-degree for workgroup-sliced top-down expansion and
  in-degree for wave-wide bottom-up probes, then the discoveries staged
  per workgroup in LDS (the level kernel's dynamic LDS, against its
  occupancy) and the hub-queue capacity;
- the multi-source search: the divisor that picks top-down levels, the
  in-degree above which a vertex's in-edges are split into wave-sized
  chunks, and the in-degree for wave-wide probes.
The thresholds never change the depths, only the time.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import graph as G
from .bfs_native import UNVISITED, BFSEngine, MultiSourceBFS

HUB_CANDIDATES = (1024, 4096, 16384)
WAVE_CANDIDATES = (8, 16, 32, 64, 128)
LDS_QUEUE_CANDIDATES = (512, 1024, 2048, 4096, 8192)
HUB_CAP_CANDIDATES = (1024, 4096, 16384)
TOPDOWN_DIVISORS = (0, 1, 2, 4, 8, 16)
MSBFS_HUB_CANDIDATES = (32, 64, 128, 256, 1024)


def union(graphs):
    """Disjoint union of `graphs` (vertex ids of graph i shifted by
    base[i]); returns (Graph, base) with base[-1] = total vertices."""
    src, dst, base = [], [], [0]
    for g in graphs:
        s, d = g.edges()
        src.append(s.astype(np.uint64) + np.uint64(base[-1]))
        dst.append(d.astype(np.uint64) + np.uint64(base[-1]))
        base.append(base[-1] + g.vertex_count)
    if base[-1] >= 2 ** 32:
        raise ValueError("the union exceeds 32-bit vertex ids")
    s = np.concatenate(src).astype(np.uint32) if src else np.zeros(0, np.uint32)
    d = np.concatenate(dst).astype(np.uint32) if dst else np.zeros(0, np.uint32)
    u = G.from_edges(base[-1], s, d)
    u.name = f"union-{len(graphs)}"
    return u, base


def small_rmat_batch(count: int = 256, seed: int = 7, log2_min: float = 10.0, log2_max: float = 16.0,
                     ef_min: int = 4, ef_max: int = 32):
    """`count` directed R-MAT graphs with vertex counts log-uniform in
    [2^log2_min, 2^log2_max] and edge factors uniform in [ef_min, ef_max]."""
    out = []
    for i in range(count):
        u1 = G.hash64(seed, 1, i) / 2.0 ** 64
        u2 = G.hash64(seed, 2, i) / 2.0 ** 64
        n = int(round(2.0 ** (log2_min + (log2_max - log2_min) * u1)))
        ef = ef_min + int(u2 * (ef_max - ef_min + 1))
        out.append(G.rmat_n(n, n * ef, seed=seed * 1000 + i))
    return out


@dataclass
class Tuned:
    params: tuple
    ms: float
    table: dict


def autotune_engine(eng: BFSEngine, roots, policy: str = "fixed", variant="td", runs: int = 2,
                    hubs=HUB_CANDIDATES, waves=WAVE_CANDIDATES, lds_queues=LDS_QUEUE_CANDIDATES,
                    hub_caps=HUB_CAP_CANDIDATES) -> Tuned:
    """On whole runs from `roots` (a root, or a list for one multi-root
    search): a grid search of (hub_degree, wave_degree) at the default
    queue sizes, then the LDS queue size, then the hub-queue capacity;
    leaves the engine set to the fastest (hub_degree, wave_degree,
    lds_queue, hub_cap)."""
    table = {}

    def time(h, w, q, c):
        if (h, w, q, c) not in table:
            eng.set_tuning(h, w)
            eng.set_queues(q, c)
            best = math.inf
            for _ in range(runs):
                best = min(best, eng.run(roots, policy, variant, download=False).total_ms)
            table[(h, w, q, c)] = best
        return table[(h, w, q, c)]

    q0, c0 = eng.queues
    h, w = min(((h, w) for h in hubs for w in waves), key=lambda hw: time(*hw, q0, c0))
    q = min(lds_queues, key=lambda x: time(h, w, x, c0))
    c = min(hub_caps, key=lambda x: time(h, w, q, x))
    params = (h, w, q, c)
    eng.set_tuning(h, w)
    eng.set_queues(q, c)
    return Tuned(params, table[params], {"/".join(map(str, k)): v for k, v in table.items()})


def autotune_msbfs(ms: MultiSourceBFS, sources, runs: int = 2, waves=WAVE_CANDIDATES + (256, 1 << 30),
                   divisors=TOPDOWN_DIVISORS, hubs=MSBFS_HUB_CANDIDATES) -> Tuned:
    """Coordinate search: the top-down divisor at the default thresholds,
    then the hub degree, then the wave threshold; returns the fastest
    (wave, divisor, hub)."""
    table = {}

    def time(w, dv, h):
        if (w, dv, h) not in table:
            ms.set_tuning(w, dv, h)
            best = math.inf
            for _ in range(runs):
                best = min(best, ms.run(sources, download=False)[1])
            table[(w, dv, h)] = best
        return table[(w, dv, h)]

    dv = min(divisors, key=lambda x: time(32, x, 128))
    h = min(hubs, key=lambda x: time(32, dv, x))
    w = min(waves, key=lambda x: time(x, dv, h))
    ms.set_tuning(w, dv, h)
    return Tuned((w, dv, h), table[(w, dv, h)], {f"{k[0]}/{k[1]}/{k[2]}": v for k, v in table.items()})


def reached_edges(deg: np.ndarray, depths: np.ndarray) -> int:
    """Out-degree sum over the reached vertices of one search."""
    return int(deg[depths != UNVISITED].sum())
