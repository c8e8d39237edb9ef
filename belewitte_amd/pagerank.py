"""PageRank implementations: the reference's 10 variants and their run.

Names follow pagerank/pagerank.cpp:205-351 of merijn/Belewitte.  Two modes:
``drop`` reproduces the reference (dangling mass lost, 100 iterations by
default, nodiv first-iteration quirk); ``uniform`` is the corrected
PageRank with the dangling mass spread uniformly.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .device import DeviceArray, DeviceGraph

DAMPENING = np.float32(0.85)   # pagerank/pagerank.hpp:9
MAX_ITERATIONS = 100           # pagerank/pagerank.cpp:55

VARIANTS = ("edge-list", "rev-edge-list", "struct-edge-list", "rev-struct-edge-list",
            "vertex-push", "vertex-pull", "vertex-push-warp", "vertex-pull-warp",
            "vertex-pull-nodiv", "vertex-pull-warp-nodiv")
REPRESENTATION = {0: "edge_list", 1: "rev_edge_list", 2: "struct_edge_list",
                  3: "rev_struct_edge_list", 4: "csr", 5: "rev_csr", 6: "csr", 7: "rev_csr",
                  8: "rev_csr", 9: "rev_csr"}
MODES = {"drop": 0, "uniform": 1}
NODIV = {8, 9}


def variant_id(name: str) -> int:
    try:
        return VARIANTS.index(name)
    except ValueError:
        names = "\n".join("    " + n for n in sorted(list(VARIANTS) + ["switch"]))
        raise KeyError(f'No kernel named "{name}" exists!\n\nSupported kernels:\n{names}') from None


def run(dg: DeviceGraph, variant: str, iterations: int = MAX_ITERATIONS, mode: str = "drop",
        warp: int = 32, chunk: int = 32, block: int = 0, ranks_buf: DeviceArray | None = None):
    """Runs pagerank/pagerank.cpp:86-127 for one variant.

    Returns (ranks float32[V], per-iteration FP64 L1 residuals)."""
    vid = variant_id(variant)
    dg.ensure(REPRESENTATION[vid])
    if vid in (4, 6) and mode == "uniform":
        dg.ensure("csr")
    V = dg.vertex_count
    own = ranks_buf is None
    ranks = DeviceArray(max(V, 1), np.float32) if own else ranks_buf
    resid = np.zeros(max(iterations, 1), dtype=np.float64)
    cfg = N.launch(block=block, warp=warp, chunk=chunk)
    N.check_hip(N.hip_lib().bw_pr_run(dg.ref, vid, iterations, MODES[mode], ranks.ptr,
                                      resid.ctypes.data_as(N.f64p), C.byref(cfg)))
    out = ranks.download()[:V]
    if own:
        ranks.free()
    return out, resid[:iterations]


def mask_bits(x: np.ndarray, n: int = 16) -> np.ndarray:
    """roundPrecision with mask_N_bits(n), pagerank/pagerank.cpp:12-35."""
    bits = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32) & np.uint32(~((1 << n) - 1) & 0xFFFFFFFF)
    return bits.view(np.float32)


def bucket_order(ranks: np.ndarray, bucket: int = 1000) -> np.ndarray:
    """Output order of pagerank/pagerank.cpp:142-168 (ids)."""
    r = mask_bits(ranks)
    ids = np.arange(r.size, dtype=np.int64)
    order = np.lexsort((ids, r))  # by rank, ties by id
    out = order.copy()
    n = out.size
    for i in range(0, n, bucket):
        end = min(i + bucket, n - 1)
        if end > i:
            out[i:end] = np.sort(out[i:end])
    return out


def format_output(ranks: np.ndarray, validate: bool = False) -> str:
    """`.output` text: ranks with 7 significant digits under --validate,
    otherwise the bucketed id order (pagerank/pagerank.cpp:134-170)."""
    if validate:
        return "".join(f"{i}\t{float(v):.7g}\n" for i, v in enumerate(ranks.tolist()))
    return "".join(f"{i}\n" for i in bucket_order(ranks).tolist())


FUSED = "vertex-pull-nodiv-fused"


class FusedPageRank:
    """The MI355X-native fused iteration of vertex-pull-nodiv: the pull
    kernel (one launch per iteration) or the binned iteration (an
    accumulate and a finalise launch per iteration, see bin()).

    Holds the plan and the workspaces, so repeated runs allocate nothing."""

    @classmethod
    def for_graph(cls, g, max_iterations: int = MAX_ITERATIONS, reorder: bool = True, on_device: bool = True):
        """Uploads `g` for the fused kernel.  reorder=True relabels the
        vertices by decreasing out-degree first (the relabelling of
        reorder-graph.cpp:40-128), which packs the most-gathered
        contributions into few cache lines: on the device from the uploaded
        graph (on_device=True), or on the host before the upload.  Ranks are
        returned in the original vertex order either way."""
        perm = None
        if reorder and on_device:
            src = DeviceGraph(g, reps=("csr", "rev_csr"))
            try:
                dg, perm = src.relabel_by_degree()
            finally:
                src.free()
            obj = cls(dg, max_iterations)
            obj.perm = perm
            obj._owns_dg = True
            return obj
        if reorder:
            g, perm = g.reorder_by_degree()
        dg = DeviceGraph(g, reps=("rev_csr",))
        obj = cls(dg, max_iterations)
        obj.perm = perm
        obj._owns_dg = True
        return obj

    def __init__(self, dg: DeviceGraph, max_iterations: int = MAX_ITERATIONS):
        if dg.offset_bytes != 4:
            raise ValueError("the fused kernel reads uint32 offsets")
        dg.ensure("rev_csr")
        self.dg = dg
        self.perm = None        # old vertex id -> id in dg, when relabelled
        self._owns_dg = False
        lib = N.hip_lib()
        p = C.c_void_p()
        N.check_hip(lib.bw_pr_plan_create(C.byref(p), dg.vertex_count, dg.edge_count,
                                          dg.desc.rev_offsets, 4, dg.desc.rev_cols))
        self._plan = p
        self.max_iterations = int(max_iterations)
        self.tuning = (0, 0, 1, 0, 16)
        self.binning = None
        V = max(dg.vertex_count, 1)
        self.ranks = DeviceArray(V, np.float32)
        self.scratch = DeviceArray(V, np.float32)
        self.acc = DeviceArray(int(lib.bw_pr_fused_acc_len(self.max_iterations)), np.float64)

    def tune(self, hot: int = 0, rows_grid: int = 0, tiles: int = 1, sorted_windows: int = 0,
             batch: int = 16) -> None:
        """Launch shape of the row phase (DROP-mode ranks never change):
        `hot` (0, 8192, 16384) most-gathered sources kept in LDS per
        workgroup, `rows_grid` persistent workgroups (0 = one per step),
        `tiles` (1, 2, 4) row tiles per workgroup and step, or
        `sorted_windows` (1, 2, 4): windows of that many tiles gathered in
        source order (0: off), `batch` (8, 16) gathers in flight per thread
        in the window kernel."""
        lib = N.hip_lib()
        N.check_hip(lib.bw_pr_plan_tune(self._plan, int(hot), int(rows_grid), int(tiles)))
        N.check_hip(lib.bw_pr_plan_sort_windows(self._plan, int(sorted_windows), int(batch)))
        self.tuning = (int(hot), int(rows_grid), int(tiles), int(sorted_windows), int(batch))

    def bin(self, rows_log2: int = 14, chunk_tasks: int = 128, lane_edges: int = 8, on_host: bool = False) -> None:
        """Switches to the binned iteration (propagation blocking): rows cut
        into partitions of 2**rows_log2 (13, 14) accumulated in LDS in FP64,
        one contribution per distinct (source, partition) pair written and
        read per iteration, partitions' edges cut into chunks of
        chunk_tasks x 2048 edges, lane_edges (8, 16) edges per lane and
        round.  rows_log2=0 returns to the pull kernels.  The plan is built
        on the device on the first call for a partition and chunk size;
        on_host=True builds it on the host instead (always anew)."""
        fn = N.hip_lib().bw_pr_plan_bin_host if on_host else N.hip_lib().bw_pr_plan_bin
        N.check_hip(fn(self._plan, int(rows_log2), int(chunk_tasks), int(lane_edges)))
        self.binning = (int(rows_log2), int(chunk_tasks), int(lane_edges)) if rows_log2 else None

    BIN_ARRAYS = {"codes": (0, np.uint16), "tasks": (1, np.uint32), "chunks": (2, np.uint32),
                  "chunk_base": (3, np.uint32), "blocks": (4, np.uint32), "run_flat": (5, np.uint32),
                  "run_bin": (6, np.uint32), "psrc": (7, np.uint16)}

    def bin_array(self, name: str) -> np.ndarray:
        """A copy of one array of the binned plan (BIN_ARRAYS; uint4 records
        come as 4 uint32 each)."""
        which, dt = self.BIN_ARRAYS[name]
        n = C.c_uint64()
        lib = N.hip_lib()
        N.check_hip(lib.bw_pr_plan_bin_array(self._plan, which, None, 0, C.byref(n)))
        out = np.empty(int(n.value) // np.dtype(dt).itemsize, dtype=dt)
        N.check_hip(lib.bw_pr_plan_bin_array(self._plan, which, out.ctypes.data_as(C.c_void_p), out.nbytes,
                                             C.byref(n)))
        return out

    def bin_info(self) -> dict:
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        N.check_hip(N.hip_lib().bw_pr_plan_bin_info(self._plan, C.byref(a), C.byref(b), C.byref(c)))
        return {"pairs": a.value, "chunks": b.value, "blocks": c.value}

    def autotune(self, iterations: int = 10, candidates=None, margin: float = 0.02) -> dict:
        """Times each launch shape over `iterations` iterations; keeps the
        first candidate (the default shape) unless another is faster by
        more than `margin`, so that run-to-run noise does not flip the
        choice.  A shape is ("bins", rows_log2, chunk_tasks, lane_edges)
        for the binned iteration, else the pull kernels' (hot, rows_grid,
        tiles, sorted_windows, batch).  Returns {shape: ms}."""
        from .device import device_props
        cus = device_props(0)["compute_units"]
        cands = candidates or [("bins", 14, 128, 8), ("bins", 13, 64, 8), ("bins", 13, 32, 8), ("bins", 14, 64, 8),
                               ("bins", 14, 80, 8), ("bins", 14, 96, 8),
                               (0, 3 * cus, 1, 2, 16), (0, 2 * cus, 1, 4, 16)]
        table = {}
        for shape in cands:
            try:
                self._apply(shape)
            except N.BelewitteError:
                continue  # a shape this graph cannot take (too many partitions)
            self.run(iterations, "uniform", download=False)
            table[shape] = min(self.run(iterations, "uniform", download=False)[2] for _ in range(3))
        best = min(table, key=table.get)
        first = next(s for s in cands if s in table)
        if table[best] > (1.0 - margin) * table[first]:
            best = first
        self._apply(best)
        return {self._shape_name(s): round(v, 4) for s, v in table.items()}

    def _apply(self, shape) -> None:
        if shape[0] == "bins":
            self.bin(*shape[1:])
        else:
            self.bin(0)
            self.tune(*shape)

    @staticmethod
    def _shape_name(shape) -> str:
        if shape[0] == "bins":
            return f"bins/rows{1 << shape[1]}/chunk{shape[2]}/lanes{shape[3]}"
        h, g, t, w, b = shape
        return f"hot{h}/grid{g}/tiles{t}" + (f"/sorted{w}/batch{b}" if w else "")

    def kernel_names(self) -> str:
        """The kernels one iteration launches under the current shape."""
        if self.binning:
            r, _, lanes = self.binning
            return f"pr_bins_accumulate_kernel<{r}, 1024, {lanes}, 2> + pr_bins_finalise_kernel<{r}, 512>"
        hot, _, tiles, width, batch = self.tuning
        rows = f"pr_windows_kernel<{width}, {batch}>" if width else f"pr_rows_kernel<{hot}, {tiles}>"
        return f"pr_pieces_kernel + {rows}"

    def info(self) -> dict:
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        N.check_hip(N.hip_lib().bw_pr_plan_info(self._plan, C.byref(a), C.byref(b), C.byref(c)))
        return {"tiles": a.value, "long_rows": b.value, "long_pieces": c.value}

    def run(self, iterations: int = 30, mode: str = "uniform", nodiv_quirk: bool = False,
            download: bool = True, stream=None, grid: int = 0):
        """Returns (ranks or None, residuals, device ms of the iteration launches)."""
        if not 1 <= iterations <= self.max_iterations:
            raise ValueError("iterations out of range of the workspace")
        resid = np.zeros(iterations, dtype=np.float64)
        ms = C.c_float(0.0)
        cfg = N.launch(grid=grid, stream=stream)
        N.check_hip(N.hip_lib().bw_pr_fused_run(
            self.dg.ref, self._plan, iterations, MODES[mode], int(nodiv_quirk), self.ranks.ptr,
            self.scratch.ptr, self.acc.ptr, resid.ctypes.data_as(N.f64p), C.byref(ms), C.byref(cfg)))
        out = None
        if download:
            out = self.ranks.download()[:self.dg.vertex_count]
            if self.perm is not None:
                out = out[self.perm]
        return out, resid, float(ms.value)

    def close(self):
        if self._plan:
            N.hip_lib().bw_pr_plan_destroy(self._plan)
            self._plan = None
        for a in (self.ranks, self.scratch, self.acc):
            a.free()
        if self._owns_dg:
            self.dg.free()
            self._owns_dg = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
