"""Small GPU workloads for counter calibration and kernel profiling.

    python -m belewitte_amd.profiling copy --mib 1024
        One streaming copy of the given size (bw_probe_copy): reads and
        writes exactly that many bytes, the reference point that scales the
        FETCH_SIZE / WRITE_SIZE counters to bytes.
    python -m belewitte_amd.profiling gather --table-mib 16 --share 1,2,4,8,16,32,64
        The gather-rate probe (bw_probe_gather): 64 Mi gathers from a
        float table where `share` consecutive lanes read one 128-byte line;
        prints lanes/s and distinct lines/s per sharing factor.
    python -m belewitte_amd.profiling pagerank --scale 22 --iterations 5 [--reorder] [--grid N --hot H --tiles T] [--sorted W --batch B] [--bins R,C,L]
        A few fused PageRank iterations on the benchmark graph (--bins: the
        binned iteration with 2^R rows per partition, chunks of C tasks, L
        edges per lane).
    python -m belewitte_amd.profiling msbfs --scale 20 --sources 64 [--wave-degree W --topdown-divisor D --hub-degree H]
        One multi-source search (64 sources) on the directed R-MAT of the
        batched BFS workload.
    python -m belewitte_amd.profiling bfs --scale 25 --sources 4 [--policy learned|td|bu|bu-wave]
        BFS engine searches on the undirected R-MAT of the learned-switch
        workload (the shipped model for the learned policy).

Run them under `rocprofv3 --pmc ... --kernel-trace -- python3 -m ...`.
The copy and gather probes live in libbw_probes.so, a profiling-only
library beside the kernel library (csrc/profiling/bw_probes.h).
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import sys

import numpy as np

_probes = None


def probes_lib() -> C.CDLL:
    """libbw_probes.so with its signatures (fails loudly when not built)."""
    global _probes
    if _probes is None:
        from . import _native as N
        path = N.LIB_DIR / "libbw_probes.so"
        if not path.exists():
            raise N.NativeLibraryMissing(f"{path} is not built; run belewitte_amd.build.build_all()")
        lib = C.CDLL(str(path), mode=os.RTLD_NOW)
        lib.bw_probe_last_error.restype = C.c_char_p
        lib.bw_probe_last_error.argtypes = []
        lib.bw_probe_copy.restype = C.c_int
        lib.bw_probe_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        lib.bw_probe_gather.restype = C.c_int
        lib.bw_probe_gather.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                        C.POINTER(C.c_float), C.c_void_p]
        _probes = lib
    return _probes


def check_probe(status: int) -> None:
    if status != 0:
        from . import _native as N
        raise N.BelewitteError(status, probes_lib().bw_probe_last_error().decode())


def copy(mib: int, repeat: int) -> None:
    from .device import DeviceArray, require_gpu, set_device, synchronize
    require_gpu()
    set_device(0)
    n = mib * (1 << 20) // 4
    src = DeviceArray(n, np.uint32)
    dst = DeviceArray(n, np.uint32)
    src.fill_u32(7)
    from .device import Event
    e0, e1 = Event(), Event()
    best = float("inf")
    for _ in range(repeat):
        e0.record()
        check_probe(probes_lib().bw_probe_copy(dst.ptr, src.ptr, n, None))
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_ms(e1))
    synchronize()
    print(f"copy: {repeat} x {n * 4} bytes, best {best:.3f} ms, {2 * n * 4 / best / 1e6:.0f} GB/s "
          f"(read + write)", flush=True)


LINE_FLOATS = 32  # floats per 128-byte cache line


def gather_indices(n: int, table_len: int, share: int, seed: int = 0, stride: int = 1) -> np.ndarray:
    """Index stream of the gather probe: groups of `share` consecutive
    entries read one random line of a `table_len`-float table; entry i
    reads word (i * stride) % 32 of it (stride 1: consecutive words,
    2: every other word, 0: all the same word)."""
    if share < 1 or share > 64 or 64 % share or table_len < LINE_FLOATS or stride < 0:
        raise ValueError("share must divide 64 and the table hold one line")
    rng = np.random.default_rng(seed)
    lines = rng.integers(0, table_len // LINE_FLOATS, size=n // share, dtype=np.int64)
    words = (np.arange(n, dtype=np.int64) * stride) % LINE_FLOATS
    return (np.repeat(lines, share) * LINE_FLOATS + words).astype(np.uint32)


def gather(table_mib: int, shares, mi: int = 64, threads: int = 256 * 8192, repeat: int = 5,
           stride: int = 1) -> dict:
    from .device import DeviceArray, require_gpu, set_device
    require_gpu()
    set_device(0)
    n = mi << 20
    threads = min(threads, n // 16)  # at least one round of 16 gathers per thread
    if threads % 256 or n % (16 * threads):
        raise ValueError("--mi must be a multiple of 32 or divide it")
    tlen = table_mib * (1 << 20) // 4
    table = DeviceArray(tlen, np.float32)
    table.upload((np.arange(tlen) % 7).astype(np.float32))
    idx = DeviceArray(n, np.uint32)
    out = DeviceArray(threads, np.float32)
    res = {}
    for share in shares:
        idx.upload(gather_indices(n, tlen, share, stride=stride))
        best = float("inf")
        for _ in range(repeat):
            ms = C.c_float()
            check_probe(probes_lib().bw_probe_gather(idx.ptr, n, table.ptr, tlen, out.ptr, threads,
                                                     C.byref(ms), None))
            best = min(best, ms.value)
        lanes = n / best / 1e6
        res[share] = {"ms": round(best, 4), "G_lanes_per_s": round(lanes, 1),
                      "G_lines_per_s": round(lanes / share, 1)}
        print(f"gather: table {table_mib} MiB, {share:2d} lanes per line, word stride {stride}: {best:.3f} ms, "
              f"{lanes:.1f} G lanes/s, {lanes / share:.1f} G lines/s", flush=True)
    for a in (table, idx, out):
        a.free()
    return res


def pagerank(scale: int, iterations: int, reorder: bool, hot: int = 0, grid: int = 0, tiles: int = 1,
             sorted_windows: int = 0, batch: int = 16, bins: str = "") -> None:
    from . import device, graph, pagerank as pr
    device.require_gpu()
    device.set_device(0)
    g = graph.rmat(scale, 16, seed=2)
    if reorder:
        g, _ = g.reorder_by_degree()
    dg = device.DeviceGraph(g, reps=("rev_csr",))
    f = pr.FusedPageRank(dg, max_iterations=iterations)
    f.tune(hot, grid, tiles, sorted_windows, batch)
    if bins:
        f.bin(*(int(x) for x in bins.split(",")))
    _, _, ms = f.run(iterations, "uniform", download=False)
    print(f"pagerank: V={g.vertex_count} E={g.edge_count} {iterations} iterations {ms:.3f} ms", flush=True)


def pagerank_variants(scale: int, iterations: int, variants) -> None:
    """The reference's PageRank variants one after another on the directed
    R-MAT (scale, edge factor 16, seed 2), uniform dangling mass: one
    warm-up run each, then one timed run (wall clock around the call, the
    download of the ranks included) for the kernel trace to attribute."""
    import time

    from . import device, graph, pagerank as pr
    device.require_gpu()
    device.set_device(0)
    g = graph.rmat(scale, 16, seed=2)
    dg = device.DeviceGraph(g, reps=device.REPRESENTATIONS)
    for v in variants:
        pr.run(dg, v, iterations, "uniform")
        device.synchronize()
        t0 = time.perf_counter()
        pr.run(dg, v, iterations, "uniform")
        dt = time.perf_counter() - t0
        print(f"pagerank {v}: V={g.vertex_count} E={g.edge_count} {iterations} iterations "
              f"{dt * 1e3:.3f} ms wall", flush=True)
    dg.free()


def msbfs(scale: int, sources: int, wave_degree: int, topdown_divisor: int, hub_degree: int) -> None:
    from . import bfs_native, device, graph
    device.require_gpu()
    device.set_device(0)
    g = graph.rmat(scale, 16, seed=6)
    dg = device.DeviceGraph(g)
    src = graph.random_sources(g, sources, seed=9)
    ms = bfs_native.MultiSourceBFS(dg)
    ms.set_tuning(wave_degree, topdown_divisor, hub_degree)
    ms.run(src, download=False)
    _, t, bits = ms.run(src, download=False)
    print(f"msbfs: V={g.vertex_count} E={g.edge_count} {len(src)} sources {t:.3f} ms, {len(bits)} levels, "
          f"new (source, vertex) pairs per level {bits.tolist()}", flush=True)
    ms.close()
    dg.free()


def bfs(scale: int, sources: int, policy: str, edge_lists: bool = True, kind: str = "rmat") -> None:
    from . import bfs_native, device, graph, switch
    device.require_gpu()
    device.set_device(0)
    if kind == "rmat":
        g = graph.rmat(scale, 16, seed=5, undirected=True)
    elif kind == "er":  # the uniform training graph: 2^24 vertices, average degree 16
        g = graph.erdos_renyi(1 << 24, 1 << 27, seed=3, undirected=True)
    elif kind == "grid":  # the grid training graph: 4096 x 4096, 1% of its edges rewired
        g = graph.grid2d(4096, 4096, rewire=0.01, seed=4)
    else:
        raise ValueError(f"unknown graph kind {kind!r}")
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg, edge_lists=edge_lists)
    if policy == "learned":
        eng.set_model(switch.Model.load(switch.MODEL_PATH).w)
    def one(s):
        if policy in ("learned", "beamer"):
            return eng.run(s, policy, "td", download=False)
        return eng.run(s, "fixed", policy, download=False)

    src = graph.random_sources(g, sources, seed=8)
    one(src[0])  # warm-up: builds the captured level batch
    total = 0.0
    for s in src:
        r = one(s)
        total += r.total_ms
        print(f"bfs {policy}: source {s} {r.num_levels} levels {r.total_ms:.3f} ms "
              f"variants {[bfs_native.VARIANTS[c] for c in r.choices]}", flush=True)
    print(f"bfs {policy}: {len(src)} sources {total:.3f} ms", flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--repeat", type=int, default=1)
    p = sub.add_parser("pagerank")
    p.add_argument("--scale", type=int, default=22)
    p.add_argument("--iterations", type=int, default=5)
    p.add_argument("--reorder", action="store_true")
    p.add_argument("--hot", type=int, default=0)
    p.add_argument("--grid", type=int, default=0)
    p.add_argument("--tiles", type=int, default=1)
    p.add_argument("--sorted", type=int, default=0, help="tiles per source-sorted window (0: off)")
    p.add_argument("--batch", type=int, default=16, help="gathers in flight per thread (window kernel)")
    p.add_argument("--bins", default="", help="binned iteration: rows_log2,chunk_tasks,lane_edges")
    q = sub.add_parser("gather")
    q.add_argument("--table-mib", type=int, default=16)
    q.add_argument("--share", default="1,2,4,8,16,32,64", help="lanes per line, comma separated")
    q.add_argument("--mi", type=int, default=64, help="gathers per launch, in Mi")
    q.add_argument("--stride", type=int, default=1, help="word stride of the lanes sharing a line")
    m = sub.add_parser("msbfs")
    m.add_argument("--scale", type=int, default=20)
    m.add_argument("--sources", type=int, default=64)
    m.add_argument("--wave-degree", type=int, default=32)
    m.add_argument("--topdown-divisor", type=int, default=4)
    m.add_argument("--hub-degree", type=int, default=128)
    b = sub.add_parser("bfs")
    b.add_argument("--scale", type=int, default=25)
    b.add_argument("--sources", type=int, default=4)
    b.add_argument("--policy", default="learned", help="learned, beamer or a variant name")
    b.add_argument("--no-edge-lists", action="store_true", help="engine without the edge-list families")
    b.add_argument("--graph", default="rmat", choices=["rmat", "er", "grid"],
                   help="undirected R-MAT of --scale, or the Erdos-Renyi or grid training graph")
    v = sub.add_parser("pagerank-variants")
    v.add_argument("--scale", type=int, default=22)
    v.add_argument("--iterations", type=int, default=10)
    v.add_argument("--variants", default="", help="comma separated (default: all ten)")
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy(a.mib, a.repeat)
    elif a.cmd == "gather":
        gather(a.table_mib, [int(x) for x in a.share.split(",")], a.mi, stride=a.stride)
    elif a.cmd == "msbfs":
        msbfs(a.scale, a.sources, a.wave_degree, a.topdown_divisor, a.hub_degree)
    elif a.cmd == "pagerank":
        pagerank(a.scale, a.iterations, a.reorder, a.hot, a.grid, a.tiles, a.sorted, a.batch, a.bins)
    elif a.cmd == "pagerank-variants":
        from .pagerank import VARIANTS
        pagerank_variants(a.scale, a.iterations, a.variants.split(",") if a.variants else VARIANTS)
    else:
        bfs(a.scale, a.sources, a.policy, not a.no_edge_lists, a.graph)
    return 0


if __name__ == "__main__":
    sys.exit(main())
