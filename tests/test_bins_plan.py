"""CPU check of the binned PageRank plan (propagation blocking).

Compiles tests/native/bins_plan_check.cpp with g++ against the plan
builder the HIP library uses, then, on small graphs: every index the two
binned kernels derive from the plan is in bounds, every vertex, pair, task
and partial slot is covered exactly once, and a sequential replay of the
kernels' arithmetic matches an FP64 PageRank of the graph within the 1e-6
parity bar.  No GPU.
"""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from belewitte_amd import graph

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("bins") / "bins_plan_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'belewitte_amd/csrc/hip'}",
                    str(ROOT / "tests/native/bins_plan_check.cpp"), "-o", str(out)], check=True)
    return out


def _write_rev_csr(g, path: Path):
    s, d = g.edges()
    order = np.lexsort((s, d))
    V = g.vertex_count
    off = np.zeros(V + 1, np.uint32)
    np.cumsum(np.bincount(d[order].astype(np.int64), minlength=V), out=off[1:])
    off.tofile(str(path) + ".off")
    s[order].astype(np.uint32).tofile(str(path) + ".cols")
    return str(path) + ".off", str(path) + ".cols"


def _hub():
    r = graph.rmat(13, 4, seed=41)
    s, d = r.edges()
    hub = np.arange(1, 6001, dtype=np.uint32)
    return graph.from_edges(r.vertex_count, np.concatenate([s, hub]),
                            np.concatenate([d, np.zeros(hub.size, np.uint32)]))


GRAPHS = {
    "rmat12": lambda: graph.rmat(12, 16, seed=31).reorder_by_degree()[0],
    "rmat14-unordered": lambda: graph.rmat(14, 8, seed=7),
    "er": lambda: graph.erdos_renyi(3000, 12000, seed=3, undirected=False),
    "star": lambda: graph.star(300),
    "chain": lambda: graph.chain(50),
    "hub": _hub,
}


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("rows_log2,chunk_tasks", [(8, 1), (10, 2), (13, 64)])
@pytest.mark.parametrize("uniform", [0, 1])
def test_plan_covers_and_matches_fp64(checker, tmp_path, name, rows_log2, chunk_tasks, uniform):
    g = GRAPHS[name]()
    off, cols = _write_rev_csr(g, tmp_path / name)
    r = subprocess.run([str(checker), off, cols, str(rows_log2), str(chunk_tasks), "20", str(uniform)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    fields = r.stdout.split()
    err = float(fields[fields.index("max_abs_err") + 1])
    ref_err = float(fields[fields.index("ref_err") + 1])
    top = float(fields[fields.index("max_rank") + 1])
    # FP64 row sums: FP32 rounding of the contributions and partials only,
    # so never less accurate than the reference's sequential FP32 sums.
    assert err <= 1e-6 and err <= 2e-6 * top and err <= ref_err + 1e-7, r.stdout


def test_plan_rejects_too_many_partitions(checker, tmp_path):
    g = graph.chain(600000)
    off, cols = _write_rev_csr(g, tmp_path / "long")
    r = subprocess.run([str(checker), off, cols, "8", "1", "1", "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 4 and "too many partitions" in r.stdout
