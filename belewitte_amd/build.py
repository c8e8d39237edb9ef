"""Build recipe of every native component (in-tree, gfx950).

    python -m belewitte_amd.build            # everything
    python -m belewitte_amd.build hip        # one component

Outputs (git-ignored, shipped with the tree):
    belewitte_amd/lib/libbwgraph.so          host graph substrate (g++ + OpenMP)
    belewitte_amd/lib/libbelewitte_hip.so    HIP kernels + C ABI (hipcc, gfx950)
    belewitte_amd/lib/libbw_probes.so        measurement probes (profiling only)
    belewitte_amd/lib/kernels/lib{bfs,pagerank}kernel.so   plugin libraries
    belewitte_amd/bin/kernel-runner          drop-in runner of the plugins
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INC = ROOT / "include"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("BW_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions stay off: some of their forms have given wrong
    # results on this target, with no error.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
    f"-I{INC}",
]

HIP_SOURCES = ["backend.hip", "bfs.hip", "pagerank.hip", "pagerank_scatter.hip", "pagerank_plan.hip", "bfs_native.hip",
               "pagerank_native.hip", "mlp.hip", "msbfs.hip"]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(map(str, cmd))}\n{out}")
    return out


def _version() -> str:
    try:
        rev = subprocess.run(["git", "rev-parse", "--short", "HEAD"], cwd=ROOT, stdout=subprocess.PIPE,
                             stderr=subprocess.DEVNULL, text=True).stdout.strip()
    except OSError:
        rev = ""
    return f"belewitte-amd-{rev or 'dev'}"


def build_graph() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libbwgraph.so"
    _run([CXX, "-std=c++17", "-O3", "-fopenmp", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{INC}",
          str(CSRC / "graph" / "graph.cpp"), "-o", str(out)])
    return out


def build_hip() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / "hip" / s for s in HIP_SOURCES if (CSRC / "hip" / s).exists()]
    ver = _version()

    def compile_one(src: Path) -> Path:
        obj = OBJ / (src.stem + ".o")
        _run([HIPCC, *HIP_FLAGS, f'-DBW_VERSION="{ver}"', "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = LIB / "libbelewitte_hip.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(out)])
    return out


def build_probes() -> Path:
    """The measurement probes of belewitte_amd.profiling (copy kernel for
    counter calibration, gather-rate probe); not part of the kernel ABI."""
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libbw_probes.so"
    _run([HIPCC, *HIP_FLAGS, "-shared", str(CSRC / "profiling" / "probes.hip"), "-o", str(out)])
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    return ROOT / "oracle" / "_build" / "liboracle.so"


def build_framework() -> list:
    """kernel-runner and the two plugin libraries (C++17 on the C ABI)."""
    fw = CSRC / "framework"
    if not (fw / "Algorithm.cpp").exists():
        return []
    (LIB / "kernels").mkdir(parents=True, exist_ok=True)
    BIN.mkdir(parents=True, exist_ok=True)
    common = [CXX, "-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", f"-I{INC}", f"-I{fw}",
              f'-DKERNEL_COMMIT="{_version()}"']
    hiplib = [f"-L{LIB}", "-lbelewitte_hip", "-lbwgraph", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN/..",
              "-Wl,-rpath,$ORIGIN/../lib"]
    fw_srcs = [str(fw / s) for s in ("Algorithm.cpp", "ImplementationBase.cpp", "Options.cpp", "Timer.cpp",
                                     "Util.cpp", "HIPBackend.cpp")]
    runner = BIN / "kernel-runner"
    outs = []

    def runner_job():
        _run(common + ["-rdynamic", str(CSRC / "runner" / "kernel-runner.cpp"), *fw_srcs, "-o", str(runner),
                       "-ldl", "-lpthread", *hiplib])
        return runner

    def plugin_job(name):
        out = LIB / "kernels" / f"lib{name}kernel.so"
        _run(common + ["-shared", str(CSRC / "plugins" / f"{name}.cpp"), "-o", str(out), *hiplib])
        return out

    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(runner_job), ex.submit(plugin_job, "bfs"), ex.submit(plugin_job, "pagerank")]
        outs = [f.result() for f in futs]
    return outs


def build_all(verbose: bool = True) -> None:
    steps = [("graph", build_graph), ("hip", build_hip), ("probes", build_probes), ("oracle", build_oracle),
             ("framework", build_framework)]
    for name, fn in steps:
        r = fn()
        if verbose:
            print(f"[build] {name}: {r}", flush=True)


def main(argv) -> int:
    targets = argv[1:] or ["all"]
    table = {"graph": build_graph, "hip": build_hip, "probes": build_probes, "oracle": build_oracle,
             "framework": build_framework, "all": build_all}
    for t in targets:
        print(f"[build] {t}: {table[t]()}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
