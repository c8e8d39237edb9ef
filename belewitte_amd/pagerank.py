"""PageRank implementations: the reference's 10 variants and their run.

Names follow pagerank/pagerank.cpp:205-351 of merijn/Belewitte.  Two modes:
``drop`` reproduces the reference (dangling mass lost, 100 iterations by
default, nodiv first-iteration quirk); ``uniform`` is the corrected
PageRank with the dangling mass spread uniformly.
"""
from __future__ import annotations

import ctypes as C
import struct

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
