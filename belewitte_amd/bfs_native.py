"""The device-driven BFS engine (MI355X-native level loop).

Computes exactly the depths and per-level frontier counts of the
reference's BFS (bfs/bfs.cpp:43-118 of merijn/Belewitte): INT_MAX for
unreached vertices, one entry per level with the number of newly
discovered vertices, the last entry 0.  Each level runs one of the engine's
variants -- the native ``td``, ``bu``, ``bu-wave`` and the reference's
eight kernel families (bfs/*.cu) on the engine's state -- picked per level
on the device by a fixed choice, a schedule, the on-device MLP, the host or
Beamer's direction-optimising rule; because every variant yields the same
depths, the choice never changes the result.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .device import DeviceGraph

NATIVE_VARIANTS = ("td", "bu", "bu-wave")
# The reference's families (bfs.KERNELS order, bfs/bfs.cpp:127-204) follow
# the native variants.
REFERENCE_VARIANTS = ("edge-list", "rev-edge-list", "struct-edge-list", "rev-struct-edge-list",
                      "vertex-push", "vertex-pull", "vertex-push-warp", "vertex-pull-warp")
VARIANTS = NATIVE_VARIANTS + REFERENCE_VARIANTS
# Representation each reference family reads beyond both CSR directions.
EDGE_LIST_REPRESENTATION = {"edge-list": "edge_list", "rev-edge-list": "rev_edge_list",
                            "struct-edge-list": "struct_edge_list", "rev-struct-edge-list": "rev_struct_edge_list"}
POLICIES = {"fixed": 0, "schedule": 1, "learned": 2, "host": 3, "beamer": 4}
NUM_FEATURES = 11
FEATURE_NAMES = ("log2_frontier_fraction", "log2_frontier_edge_fraction", "log2_unvisited_fraction", "depth",
                 "log2_mean_degree", "log2_std_degree", "prev_variant", "log2_prev_time_ns", "log2_vertices",
                 "log2_unexplored_edge_fraction", "log2_frontier_growth")
UNVISITED = 2147483647
DEFAULT_HUB_DEGREE = 4096
DEFAULT_WAVE_DEGREE = 8
DEFAULT_LDS_QUEUE = 2048
DEFAULT_HUB_CAP = 4096


def variant_id(v) -> int:
    if isinstance(v, (int, np.integer)):
        if not 0 <= int(v) < len(VARIANTS):
            raise KeyError(f"unknown native BFS variant {v}")
        return int(v)
    try:
        return VARIANTS.index(v)
    except ValueError:
        raise KeyError(f"unknown native BFS variant {v!r}; known: {', '.join(VARIANTS)}") from None


@dataclass
class EngineResult:
    levels: np.ndarray | None   # int32 depth per vertex (None if not downloaded)
    counts: np.ndarray          # newly discovered vertices per level (last is 0)
    times_ns: np.ndarray        # device time per level
    choices: np.ndarray         # variant per level
    total_ms: float             # device time of the whole search

    @property
    def num_levels(self) -> int:
        return int(self.counts.size)

    @property
    def reached(self) -> int:
        return int(self.counts.sum()) + 1


@dataclass
class LevelRecord:
    """One level of an exhaustive run: its features and every variant's time."""
    source: int
    depth: int
    features: np.ndarray        # float32[NUM_FEATURES]
    ms: np.ndarray              # float64[K] device time of each variant's level kernel
    count: int                  # vertices discovered at this level


class BFSEngine:
    """Holds the engine's device state for one graph (both CSR directions,
    uint32 offsets).  Repeated searches allocate nothing.

    The edge-list families run when their representation is resident when
    the engine is made: edge_lists=True uploads all four; otherwise those
    the DeviceGraph already holds count."""

    def __init__(self, dg: DeviceGraph, max_levels: int | None = None, edge_lists: bool = False):
        if dg.offset_bytes != 4:
            raise ValueError("the BFS engine reads uint32 offsets")
        dg.ensure("csr")
        dg.ensure("rev_csr")
        if edge_lists:
            for rep in EDGE_LIST_REPRESENTATION.values():
                dg.ensure(rep)
        self.dg = dg
        self.available = tuple(v for v in VARIANTS
                               if v not in EDGE_LIST_REPRESENTATION or dg.edge_count == 0
                               or dg.has(EDGE_LIST_REPRESENTATION[v]))
        V = dg.vertex_count
        self.max_levels = int(max_levels or min(V + 1, 1 << 22))
        p = C.c_void_p()
        N.check_hip(N.hip_lib().bw_bfs_engine_create(C.byref(p), dg.ref, None, self.max_levels))
        self._e = p
        self._counts = np.zeros(self.max_levels, dtype=np.uint32)
        self._times = np.zeros(self.max_levels, dtype=np.uint64)
        self._choices = np.zeros(self.max_levels, dtype=np.int32)
        self.model_outputs = 0
        self.tuning = (DEFAULT_HUB_DEGREE, DEFAULT_WAVE_DEGREE)
        self.queues = (DEFAULT_LDS_QUEUE, DEFAULT_HUB_CAP)

    # ------------------------------------------------------------ set-up

    def set_model(self, weights: np.ndarray) -> None:
        """A packed model with one output per native variant or per engine
        variant (its size tells which)."""
        w = np.ascontiguousarray(weights, dtype=np.float32)
        lib = N.hip_lib()
        k = next((k for k in (len(VARIANTS), len(NATIVE_VARIANTS)) if w.size == int(lib.bw_bfs_engine_mlp_len(k))),
                 None)
        if k is None:
            raise ValueError("model size does not match the engine's variant count")
        N.check_hip(N.hip_lib().bw_bfs_engine_set_model(self._e, w.ctypes.data_as(N.f32p), k))
        self.model_outputs = k

    def set_schedule(self, schedule) -> int:
        s = np.ascontiguousarray([variant_id(v) for v in schedule], dtype=np.int32)
        N.check_hip(N.hip_lib().bw_bfs_engine_set_schedule(self._e, s.ctypes.data_as(N.i32p), s.size))
        return int(s.size)

    def set_tuning(self, hub_degree: int = DEFAULT_HUB_DEGREE, wave_degree: int = DEFAULT_WAVE_DEGREE) -> None:
        N.check_hip(N.hip_lib().bw_bfs_engine_set_tuning(self._e, int(hub_degree), int(wave_degree)))
        self.tuning = (int(hub_degree), int(wave_degree))

    def set_queues(self, lds_queue: int = DEFAULT_LDS_QUEUE, hub_cap: int = DEFAULT_HUB_CAP) -> None:
        """Discoveries staged per workgroup in LDS and hub-queue capacity
        (bw_bfs_engine_set_queues); depths never depend on either."""
        N.check_hip(N.hip_lib().bw_bfs_engine_set_queues(self._e, int(lds_queue), int(hub_cap)))
        self.queues = (int(lds_queue), int(hub_cap))

    # -------------------------------------------------------------- runs

    def run(self, root, policy: str = "fixed", variant="td", schedule=None, levels_per_sync: int = 8,
            use_graph: bool = True, download: bool = True, stream=None) -> EngineResult:
        """A whole search with the level loop on the device.  `root` may be
        a list of roots, all at depth 0 (a multi-root search)."""
        pol = POLICIES[policy]
        slen = self.set_schedule(schedule) if schedule is not None else 1
        if policy == "schedule" and schedule is None:
            raise ValueError("the schedule policy needs a schedule")
        n = C.c_uint32(0)
        ms = C.c_float(0.0)
        outs = (self._counts.ctypes.data_as(N.u32p), self._times.ctypes.data_as(N.u64p),
                self._choices.ctypes.data_as(N.i32p), C.byref(n), C.byref(ms), stream)
        if np.ndim(root) == 0:
            N.check_hip(N.hip_lib().bw_bfs_engine_run(
                self._e, int(root), pol, variant_id(variant), slen, int(levels_per_sync), int(bool(use_graph)),
                *outs))
        else:
            roots = np.ascontiguousarray(root, dtype=np.uint32)
            N.check_hip(N.hip_lib().bw_bfs_engine_run_multi(
                self._e, roots.ctypes.data_as(N.u32p), roots.size, pol, variant_id(variant), slen,
                int(levels_per_sync), int(bool(use_graph)), *outs))
        k = min(int(n.value), self.max_levels)
        return EngineResult(self.levels() if download else None, self._counts[:k].copy(),
                            self._times[:k].copy(), self._choices[:k].copy(), float(ms.value))

    def levels(self) -> np.ndarray:
        p = C.c_void_p()
        N.check_hip(N.hip_lib().bw_bfs_engine_levels(self._e, C.byref(p)))
        V = self.dg.vertex_count
        out = np.empty(V, dtype=np.int32)
        N.check_hip(N.hip_lib().bw_hip_memcpy_d2h(out.ctypes.data_as(C.c_void_p), p, 4 * V))
        return out

    def status(self, stream=None) -> dict:
        done, depth, nl = C.c_uint32(), C.c_uint32(), C.c_uint32()
        choice = C.c_int32()
        feat = np.zeros(NUM_FEATURES, dtype=np.float32)
        N.check_hip(N.hip_lib().bw_bfs_engine_status(self._e, C.byref(done), C.byref(depth), C.byref(nl),
                                                     C.byref(choice), feat.ctypes.data_as(N.f32p), stream))
        return {"done": bool(done.value), "depth": int(depth.value), "num_levels": int(nl.value),
                "choice": int(choice.value), "features": feat}

    def exhaustive(self, root: int, repeats: int = 1, stream=None, variants=None):
        """Times every available variant (or those named) at every level
        from the same saved state; the others' times are inf.

        The search continues along the per-level fastest variant, so the
        features "previous variant" and "previous time" are those of the
        per-level oracle.  Returns (levels, counts, [LevelRecord])."""
        lib = N.hip_lib()
        K = len(VARIANTS)
        run = [variant_id(v) for v in (variants if variants is not None else self.available)]
        for k in run:
            if VARIANTS[k] not in self.available:
                raise ValueError(f"variant {VARIANTS[k]} needs a representation the engine was not given")
        N.check_hip(lib.bw_bfs_engine_begin(self._e, int(root), POLICIES["host"], 0, 1, stream))
        records, counts = [], []
        ms, cnt, done = C.c_float(), C.c_uint32(), C.c_uint32()
        while True:
            st = self.status(stream)
            if st["done"]:
                break
            if st["depth"] + 1 >= self.max_levels:
                raise N.BelewitteError(5, "exhaustive BFS reached max_levels")
            N.check_hip(lib.bw_bfs_engine_snapshot(self._e, 0, stream))
            times = np.full(K, np.inf)
            seen = set()
            for r in range(repeats):
                for k in run:
                    N.check_hip(lib.bw_bfs_engine_snapshot(self._e, 1, stream))
                    N.check_hip(lib.bw_bfs_engine_step_forced(self._e, k, C.byref(ms), C.byref(cnt),
                                                              C.byref(done), stream))
                    times[k] = min(times[k], float(ms.value))
                    seen.add(int(cnt.value))
            if len(seen) != 1:
                raise N.BelewitteError(2, f"variants disagree on the level-{st['depth']} count: {sorted(seen)}")
            best = int(np.argmin(times))
            N.check_hip(lib.bw_bfs_engine_snapshot(self._e, 1, stream))
            N.check_hip(lib.bw_bfs_engine_step_forced(self._e, best, C.byref(ms), C.byref(cnt),
                                                      C.byref(done), stream))
            records.append(LevelRecord(int(root), st["depth"], st["features"].copy(), times, int(cnt.value)))
            counts.append(int(cnt.value))
            if done.value:
                break
        return self.levels(), np.asarray(counts, dtype=np.uint32), records

    def close(self) -> None:
        if self._e:
            N.hip_lib().bw_bfs_engine_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiSourceBFS:
    """Up to 64 searches on one graph sharing every pass over it
    (bw_msbfs_*): depths[i] is the BFS depth array from sources[i]."""

    MAX_SOURCES = 64

    def __init__(self, dg: DeviceGraph, max_levels: int | None = None):
        if dg.offset_bytes != 4:
            raise ValueError("multi-source BFS reads uint32 offsets")
        dg.ensure("rev_csr")
        dg.ensure("csr")
        self.dg = dg
        self.max_levels = int(max_levels or min(dg.vertex_count + 1, 1 << 22))
        p = C.c_void_p()
        N.check_hip(N.hip_lib().bw_msbfs_create(C.byref(p), dg.ref, self.max_levels))
        self._m = p
        self._depths = None
        self._bits = np.zeros(self.max_levels, dtype=np.uint64)
        self.wave_degree = 32
        self.topdown_divisor = 4
        self.hub_degree = 128

    def set_tuning(self, wave_degree: int = 32, topdown_divisor: int = 4, hub_degree: int = 128) -> None:
        """wave_degree: in-degree above which a whole wave probes a vertex
        (bottom-up levels); hub_degree: in-degree above which a vertex's
        in-edges are split into chunks of 512, one wave each (bottom-up);
        a level runs top-down when its frontier's out-edges times
        topdown_divisor are at most E (0: never).  Depths depend on none."""
        N.check_hip(N.hip_lib().bw_msbfs_set_tuning(self._m, int(wave_degree), int(topdown_divisor),
                                                    int(hub_degree)))
        self.wave_degree = int(wave_degree)
        self.topdown_divisor = int(topdown_divisor)
        self.hub_degree = int(hub_degree)

    def run(self, sources, levels_per_sync: int = 8, use_graph: bool = True, download: bool = True, stream=None):
        """Returns (depths int32[n][V] or None, total_ms, level_bits)."""
        from .device import DeviceArray
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        n, V = src.size, self.dg.vertex_count
        if not 1 <= n <= self.MAX_SOURCES:
            raise ValueError("multi-source BFS takes 1..64 sources")
        if self._depths is None or self._depths.count < n * V:
            if self._depths is not None:
                self._depths.free()
            self._depths = DeviceArray(self.MAX_SOURCES * V, np.int32)
        nl, ms = C.c_uint32(), C.c_float()
        N.check_hip(N.hip_lib().bw_msbfs_run(self._m, src.ctypes.data_as(N.u32p), n, self._depths.ptr,
                                             int(levels_per_sync), int(bool(use_graph)),
                                             self._bits.ctypes.data_as(N.u64p), C.byref(nl), C.byref(ms), stream))
        depths = None
        if download:
            flat = np.empty(n * V, dtype=np.int32)
            N.check_hip(N.hip_lib().bw_hip_memcpy_d2h(flat.ctypes.data_as(C.c_void_p), self._depths.ptr, 4 * n * V))
            depths = flat.reshape(n, V)
        return depths, float(ms.value), self._bits[:min(int(nl.value), self.max_levels)].copy()

    def close(self) -> None:
        if self._m:
            N.hip_lib().bw_msbfs_destroy(self._m)
            self._m = None
        if self._depths is not None:
            self._depths.free()
            self._depths = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def teps_edges(g, levels: np.ndarray) -> int:
    """Edges counted by the TEPS figures: the out-degree sum over reached
    vertices (directed), halved for undirected graphs (Graph500)."""
    deg = np.diff(g.offsets()).astype(np.int64)
    e = int(deg[levels != UNVISITED].sum())
    return e // 2 if g.undirected else e
