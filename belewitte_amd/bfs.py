"""BFS implementations: the reference's 32 names and the level loop.

Names and kernel families follow bfs/bfs.cpp:121-224 of merijn/Belewitte:
eight kernels (edge-list, rev-edge-list, struct-edge-list,
rev-struct-edge-list, vertex-push, vertex-pull, vertex-push-warp,
vertex-pull-warp) times four frontier reductions (suffixes "", "-bulk",
"-warpreduce", "-blockreduce").  Every implementation computes the same
depths, so any per-level schedule of them does too.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .device import DeviceArray, DeviceGraph

UNVISITED = 2147483647  # INT_MAX, bfs/bfs.cpp:71-72

KERNELS = ("edge-list", "rev-edge-list", "struct-edge-list", "rev-struct-edge-list",
           "vertex-push", "vertex-pull", "vertex-push-warp", "vertex-pull-warp")
REDUCTIONS = {"": 0, "-bulk": 1, "-warpreduce": 2, "-blockreduce": 3}

# Representation each kernel family reads (bfs/bfs.cpp:127-204 tags).
REPRESENTATION = {0: "edge_list", 1: "rev_edge_list", 2: "struct_edge_list",
                  3: "rev_struct_edge_list", 4: "csr", 5: "rev_csr", 6: "csr", 7: "rev_csr"}
WARP_KERNELS = {6, 7}


def implementations() -> dict:
    """name -> (kernel id, reduction id), the 32 names of bfs/bfs.cpp."""
    out = {}
    for suffix, red in REDUCTIONS.items():
        for k, name in enumerate(KERNELS):
            out[name + suffix] = (k, red)
    return dict(sorted(out.items()))


IMPLEMENTATIONS = implementations()


def parse(name: str):
    try:
        return IMPLEMENTATIONS[name]
    except KeyError:
        names = "\n".join("    " + n for n in sorted(list(IMPLEMENTATIONS) + ["switch"]))
        raise KeyError(f'No kernel named "{name}" exists!\n\nSupported kernels:\n{names}') from None


@dataclass
class BFSResult:
    levels: np.ndarray          # int32 depth per vertex, UNVISITED if unreached
    counts: np.ndarray          # frontier count of each launch
    level_ms: np.ndarray        # device time of each launch
    schedule: list = field(default_factory=list)

    @property
    def num_levels(self) -> int:
        return int(self.counts.size)

    def reached_edges(self, g) -> int:
        """Edges scanned from reached vertices (E_r of the TEPS metric)."""
        deg = np.diff(g.offsets()).astype(np.int64)
        return int(deg[self.levels != UNVISITED].sum())


def run(dg: DeviceGraph, root: int, implementation="vertex-push", schedule=None,
        warp: int = 32, chunk: int = 32, block: int = 0, max_levels: int | None = None,
        levels_buf: DeviceArray | None = None, time_levels: bool = True) -> BFSResult:
    """Runs the level loop of bfs/bfs.cpp:67-107.

    `schedule` (a list of implementation names) overrides `implementation`
    and picks schedule[level % len] per level; all entries must share one
    reduction suffix only for bookkeeping (each launch uses its own)."""
    names = list(schedule) if schedule else [implementation]
    parsed = [parse(n) for n in names]
    reductions = {r for _, r in parsed}
    kernel_ids = [k for k, _ in parsed]
    for k in kernel_ids:
        dg.ensure(REPRESENTATION[k])
    V = dg.vertex_count
    if max_levels is None:
        max_levels = V + 2
    own = levels_buf is None
    levels = DeviceArray(max(V, 1), np.int32) if own else levels_buf
    counts = np.zeros(max(max_levels, 1), dtype=np.uint32)
    ms = np.zeros(max(max_levels, 1), dtype=np.float32)
    nlev = C.c_int64(0)
    cfg = N.launch(block=block, warp=warp, chunk=chunk)
    lib = N.hip_lib()
    if len(reductions) == 1:
        sched = (C.c_int32 * len(kernel_ids))(*kernel_ids)
        N.check_hip(lib.bw_bfs_run(dg.ref, sched, len(kernel_ids), reductions.pop(), root, levels.ptr,
                                   counts.ctypes.data_as(N.u32p),
                                   ms.ctypes.data_as(N.f32p) if time_levels else None,
                                   max_levels, C.byref(nlev), C.byref(cfg)))
        n = int(nlev.value)
    else:
        n = _run_mixed(dg, root, parsed, levels, counts, ms, max_levels, cfg)
    host = levels.download()[:V] if V else np.zeros(0, np.int32)
    if own:
        levels.free()
    if root >= V:
        host = np.zeros(0, np.int32)
    return BFSResult(host, counts[:n].copy(), ms[:n].copy(), names)


def _run_mixed(dg, root, parsed, levels, counts, ms, max_levels, cfg) -> int:
    """Level loop over implementations with different reductions: one
    bw_bfs_level launch per level, frontier read back after each."""
    from .device import synchronize
    lib = N.hip_lib()
    V = dg.vertex_count
    if root >= V:
        return 0
    init = np.full(V, UNVISITED, dtype=np.int32)
    init[root] = 0
    levels.upload(init)
    frontier = DeviceArray(1, np.uint32)
    cnt = np.zeros(1, np.uint32)
    depth = 0
    while True:
        if depth >= max_levels:
            raise N.BelewitteError(5, "BFS did not finish within max_levels launches")
        frontier.fill_u32(0)
        k, r = parsed[depth % len(parsed)]
        N.check_hip(lib.bw_bfs_level(dg.ref, k, r, levels.ptr, depth, frontier.ptr, C.byref(cfg)))
        synchronize()
        frontier.download(cnt)
        counts[depth] = cnt[0]
        depth += 1
        if cnt[0] == 0:
            break
    frontier.free()
    return depth


def format_output(levels: np.ndarray) -> str:
    """`.output` text of bfs/bfs.cpp:115-117 ("i\\t<depth>" per vertex, C locale)."""
    if levels.size == 0:
        return ""
    idx = np.arange(levels.size)
    return "".join(f"{i}\t{d}\n" for i, d in zip(idx.tolist(), levels.tolist()))
