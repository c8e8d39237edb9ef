"""The drop-in path on the GPU: kernel-runner + lib{bfs,pagerank}kernel.so.

Runs the reference's command lines (kernel-runner.cpp:62-89, the -S stdin
batch protocol of benchmark-analysis/ingest-src/JobPool.hs:56-73) against a
`.graph` file and checks the `.output` files against the oracle, the
`.timings` grammar and the `commit:label` stdout protocol.  The switch is
exercised with a predictor library of the reference's ABI
(ImplementationTemplate.hpp:767-827).
"""
import os
import re
import subprocess
import textwrap
from pathlib import Path

import numpy as np
import pytest

from belewitte_amd import bfs, graph, pagerank
from oracle import binding as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
RUNNER = ROOT / "belewitte_amd" / "bin" / "kernel-runner"
TIMING = re.compile(r"^[\w:]+:[-+0-9.e]+ [-+0-9.e]+ [-+0-9.e]+ [-+0-9.e]+$")


def run(args, stdin=None, cwd=None):
    env = dict(os.environ, LANG="C", LC_ALL="C")
    r = subprocess.run([str(RUNNER), *args], input=stdin, capture_output=True, text=True, cwd=cwd, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    return r


@pytest.fixture(scope="module")
def gfile(tmp_path_factory, gpu):
    d = tmp_path_factory.mktemp("g")
    g = graph.rmat(11, 8, seed=17)
    p = d / "rmat11.graph"
    g.save(p)
    return g, p


def parse_bfs(text):
    rows = [l.split("\t") for l in text.splitlines()]
    assert all(len(r) == 2 for r in rows)
    return np.array([int(r[1]) for r in rows], dtype=np.int64)


def test_bfs_batch_all_implementations(gfile, tmp_path):
    g, p = gfile
    root = g.max_out_degree_vertex()
    want, nlev = O.bfs_queue(O.CSR.from_graph(g), root)
    names = sorted(bfs.IMPLEMENTATIONS) + ["switch"]
    jobs = "".join(f'"{n}" -a bfs -k {n} -r {root} {p}\n' for n in names)
    r = run(["-S", "-o", str(tmp_path)], stdin=jobs)
    lines = r.stdout.split()
    assert [l.split(":", 1)[1] for l in lines] == names
    assert all(l.startswith("belewitte-amd") for l in lines)
    for n in names:
        got = parse_bfs((tmp_path / f"{n}.output").read_text())
        assert np.array_equal(got, want.astype(np.int64)), n
        timings = (tmp_path / f"{n}.timings").read_text().splitlines()
        assert all(TIMING.match(t) for t in timings), timings
        keys = {t.split(":")[0] if not t[0].isdigit() else ":".join(t.split(":")[:2]) for t in timings}
        assert {"graphTransfer", "initResults", "computation", "resultTransfer"} <= keys
        assert sum(1 for k in keys if k.endswith(":bfsLevel")) == nlev


def test_bfs_root_out_of_range_gives_empty_output(gfile, tmp_path):
    g, p = gfile
    run(["-a", "bfs", "-k", "vertex-pull", "-r", str(g.vertex_count + 5), "-o", str(tmp_path), str(p)])
    assert (tmp_path / "rmat11.output").read_text() == ""


def test_pagerank_validate_and_bucket_output(gfile, tmp_path):
    g, p = gfile
    csr = O.CSR.from_graph(g)
    for name in pagerank.VARIANTS:
        run(["-a", "pagerank", "-k", name, "--validate", "-o", str(tmp_path), str(p)])
        vals = np.array([float(l.split("\t")[1]) for l in (tmp_path / "rmat11.output").read_text().splitlines()])
        want, _ = O.pagerank_f32(csr, pagerank.variant_id(name), 100, "drop")
        assert np.max(np.abs(vals - want)) <= 1e-6 + 1e-6 * np.abs(want).max(), name
    # Default output: the bucketed order, exact for the deterministic pull.
    run(["-a", "pagerank", "-k", "vertex-pull", "-o", str(tmp_path), str(p)])
    order = [int(l) for l in (tmp_path / "rmat11.output").read_text().split()]
    want, _ = O.pagerank_f32(csr, pagerank.variant_id("vertex-pull"), 100, "drop")
    assert order == O.bucket_order(want).astype(int).tolist()


def test_pagerank_uniform_iterations_flags(gfile, tmp_path):
    g, p = gfile
    csr = O.CSR.from_graph(g)
    for name in ("vertex-push", "vertex-pull-nodiv", "edge-list"):
        run(["-a", "pagerank", "-k", name, "--validate", "--dangling", "uniform", "--iterations", "30", "-o",
             str(tmp_path), str(p)])
        vals = np.array([float(l.split("\t")[1]) for l in (tmp_path / "rmat11.output").read_text().splitlines()])
        want, _ = O.pagerank_f32(csr, pagerank.variant_id(name), 30, "uniform")
        assert np.max(np.abs(vals - want)) <= 2e-6, name


PREDICTOR = textwrap.dedent("""
    #include <cstdint>
    #include <functional>
    #include <map>
    #include <string>
    #include <tuple>
    #include <vector>
    static double frontier = 0, visited = 0;
    using Impls = std::vector<std::tuple<std::string,int64_t,size_t,size_t,size_t>>;
    using Props = std::map<std::string, std::reference_wrapper<double>>;
    // (name, id, index, warp, chunk)
    extern "C" const Impls implNames;
    extern "C" const Impls implNames = {
        {"vertex-push-warpreduce", 11, 0, 32, 32}, {"vertex-pull-bulk", 12, 1, 32, 32},
        {"edge-list-blockreduce", 13, 2, 32, 32}, {"vertex-pull-warp", 14, 3, 16, 64}};
    extern "C" const Props propNames;
    extern "C" const Props propNames = {{"frontier abs", std::ref(frontier)}, {"visited abs", std::ref(visited)}};
    extern "C" const int32_t default_impl;
    extern "C" const int32_t default_impl = 0;
    extern "C" int32_t lookup();
    extern "C" int32_t lookup() { return static_cast<int32_t>(static_cast<int64_t>(frontier + visited) % 4); }
""")


def test_switch_with_predictor_and_log(gfile, tmp_path):
    g, p = gfile
    src = tmp_path / "pred.cpp"
    src.write_text(PREDICTOR)
    so = tmp_path / "libpred.so"
    subprocess.run(["g++", "-std=c++17", "-shared", "-fPIC", str(src), "-o", str(so)], check=True)
    root = g.max_out_degree_vertex()
    want, nlev = O.bfs_queue(O.CSR.from_graph(g), root)
    log = tmp_path / "props.log"
    run(["-a", "bfs", "-k", "switch", "-r", str(root), "-m", str(so), "-l", str(log), "-o", str(tmp_path), str(p)])
    assert np.array_equal(parse_bfs((tmp_path / "rmat11.output").read_text()), want.astype(np.int64))
    lines = log.read_text().splitlines()
    assert any(l.startswith("graph:vertex count:") for l in lines)
    preds = [l for l in lines if l.startswith("prediction:")]
    assert len(preds) == nlev
    assert {int(l.split(":")[2]) for l in preds} <= {11, 12, 13, 14}
    steps = [l for l in lines if l.startswith("step:")]
    assert any(":frontier abs:" in l for l in steps) and any(":visited rel:" in l for l in steps)
    # Without a model the switch runs edge-list.
    run(["-a", "bfs", "-k", "switch", "-r", str(root), "-o", str(tmp_path), str(p)])
    assert np.array_equal(parse_bfs((tmp_path / "rmat11.output").read_text()), want.astype(np.int64))


def test_native_engine_through_the_runner(gfile, tmp_path):
    """The device-driven engine as kernel-runner implementations: fixed
    variants (native and reference families), a schedule, Beamer's rule
    and the learned switch with an 11-output and a 3-output model file."""
    from belewitte_amd import bfs_native, switch
    g, p = gfile
    root = g.max_out_degree_vertex()
    want, nlev = O.bfs_queue(O.CSR.from_graph(g), root)
    X = np.random.default_rng(0).normal(8, 3, (32, bfs_native.NUM_FEATURES))
    model, model3 = tmp_path / "m.json", tmp_path / "m3.json"
    switch.Model.init(X, seed=1, target_mean=9.0).save(model)
    switch.Model.init(X, outputs=3, seed=2, target_mean=9.0).save(model3)
    jobs = [("native", ["--variant", "td"]), ("native", ["--variant", "bu-wave"]),
            ("native", ["--variant", "edge-list"]), ("native", ["--variant", "vertex-pull-warp"]),
            ("native-schedule", ["--schedule", "bu,td,rev-struct-edge-list,bu-wave"]),
            ("native-beamer", []),
            ("native-learned", ["--model", str(model)]), ("native-learned", ["--model", str(model3)]),
            ("native-learned", [])]
    for i, (name, extra) in enumerate(jobs):
        out = tmp_path / str(i)
        out.mkdir()
        run(["-a", "bfs", "-k", name, "-r", str(root), *extra, "-o", str(out), str(p)])
        got = parse_bfs((out / "rmat11.output").read_text())
        assert np.array_equal(got, want.astype(np.int64)), (name, extra)
        timings = (out / "rmat11.timings").read_text().splitlines()
        assert all(TIMING.match(t) for t in timings), timings
        assert sum(1 for t in timings if ":bfsLevel:" in t) == nlev


def test_fused_pagerank_through_the_runner(gfile, tmp_path):
    """vertex-pull-nodiv-fused computes vertex-pull-nodiv: DROP mode with the
    reference's undivided first iteration; its sums (compensated in the pull
    kernels, FP64 in the binned iteration, the default) are never less
    accurate than the reference's sequential FP32 ones, so the bar is 1e-6
    plus that path's own error against the FP64 recurrence."""
    g, p = gfile
    csr = O.CSR.from_graph(g)
    vid = pagerank.variant_id("vertex-pull-nodiv")
    cases = [([], 100, "drop"), (["--dangling", "uniform", "--iterations", "30"], 30, "uniform"),
             (["--no-reorder", "--bins", "0", "--windows", "0"], 100, "drop"),
             (["--bins", "0", "--windows", "4", "--rows-grid", "7"], 100, "drop"),
             (["--bins", "13", "--bin-chunk", "1"], 100, "drop")]
    for extra, iters, mode in cases:
        run(["-a", "pagerank", "-k", "vertex-pull-nodiv-fused", "--validate", *extra, "-o", str(tmp_path), str(p)])
        vals = np.array([float(l.split("\t")[1]) for l in (tmp_path / "rmat11.output").read_text().splitlines()])
        want, _ = O.pagerank_f32(csr, vid, iters, mode)
        x = O.pagerank_f64(csr, iters, mode, nodiv_quirk=(mode == "drop"))
        ref_err = float(np.max(np.abs(want.astype(np.float64) - x)))
        assert vals.size == g.vertex_count
        assert float(np.max(np.abs(vals - want))) <= 1e-6 + ref_err, extra
    # Default output: the bucketed order of the same ranks.
    run(["-a", "pagerank", "-k", "vertex-pull-nodiv-fused", "-o", str(tmp_path), str(p)])
    order = [int(l) for l in (tmp_path / "rmat11.output").read_text().split()]
    assert sorted(order) == list(range(g.vertex_count))
    timings = (tmp_path / "rmat11.timings").read_text().splitlines() if (tmp_path / "rmat11.timings").exists() else []
    assert all(TIMING.match(t) for t in timings)


def test_fused_matches_pull_nodiv_on_duplicate_edges(tmp_path, gpu):
    """A file with duplicate edges: the fused implementation (relabelled by
    degree on load, binned by default) ranks the same multigraph as
    vertex-pull-nodiv, within 1e-6 plus the reference path's own error."""
    from graphfiles import duplicate_edge_graph
    p = tmp_path / "dup.graph"
    duplicate_edge_graph(p)
    g = graph.load(p)
    csr = O.CSR.from_graph(g)
    out = {}
    for k in ("vertex-pull-nodiv", "vertex-pull-nodiv-fused"):
        run(["-a", "pagerank", "-k", k, "--validate", "-o", str(tmp_path), str(p)])
        out[k] = np.array([float(l.split("\t")[1]) for l in (tmp_path / "dup.output").read_text().splitlines()])
    want, _ = O.pagerank_f32(csr, pagerank.variant_id("vertex-pull-nodiv"), 100, "drop")
    x = O.pagerank_f64(csr, 100, "drop", nodiv_quirk=True)
    ref_err = float(np.max(np.abs(want.astype(np.float64) - x)))
    assert out["vertex-pull-nodiv"].size == g.vertex_count == out["vertex-pull-nodiv-fused"].size
    assert float(np.max(np.abs(out["vertex-pull-nodiv-fused"] - out["vertex-pull-nodiv"]))) <= 1e-6 + ref_err
