"""Small GPU workloads for counter calibration and kernel profiling.

    python -m belewitte_amd.profiling copy --mib 1024
        One streaming copy of the given size (bw_hip_copy_kernel): reads and
        writes exactly that many bytes, the reference point that scales the
        FETCH_SIZE / WRITE_SIZE counters to bytes.
    python -m belewitte_amd.profiling pagerank --scale 22 --iterations 5
        A few fused PageRank iterations on the benchmark graph.

Run them under `rocprofv3 --pmc ... --kernel-trace -- python3 -m ...`.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def copy(mib: int, repeat: int) -> None:
    from . import _native as N
    from .device import DeviceArray, require_gpu, set_device, synchronize
    require_gpu()
    set_device(0)
    n = mib * (1 << 20) // 4
    src = DeviceArray(n, np.uint32)
    dst = DeviceArray(n, np.uint32)
    src.fill_u32(7)
    for _ in range(repeat):
        N.check_hip(N.hip_lib().bw_hip_copy_kernel(dst.ptr, src.ptr, n, None))
    synchronize()
    print(f"copy: {repeat} x {n * 4} bytes", flush=True)


def pagerank(scale: int, iterations: int, reorder: bool) -> None:
    from . import device, graph, pagerank as pr
    device.require_gpu()
    device.set_device(0)
    g = graph.rmat(scale, 16, seed=2)
    if reorder:
        g, _ = g.reorder_by_degree()
    dg = device.DeviceGraph(g, reps=("rev_csr",))
    f = pr.FusedPageRank(dg, max_iterations=iterations)
    _, _, ms = f.run(iterations, "uniform", download=False)
    print(f"pagerank: V={g.vertex_count} E={g.edge_count} {iterations} iterations {ms:.3f} ms", flush=True)


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
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        copy(a.mib, a.repeat)
    else:
        pagerank(a.scale, a.iterations, a.reorder)
    return 0


if __name__ == "__main__":
    sys.exit(main())
