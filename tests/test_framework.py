"""kernel-runner and the plugin libraries on the CPU (no kernel launches).

Covers plugin discovery (lib<algo>kernel.so, registerHIP), the registered
implementation names of bfs/bfs.cpp:121-224 and pagerank/pagerank.cpp:174-358,
the version query, per-implementation flags, and the error paths a CPU can
force (unknown algorithm, no device).
"""
import os
import subprocess
from pathlib import Path

import pytest

from belewitte_amd import bfs, pagerank

ROOT = Path(__file__).resolve().parents[1]
RUNNER = ROOT / "belewitte_amd" / "bin" / "kernel-runner"

pytestmark = pytest.mark.skipif(not RUNNER.exists(), reason="kernel-runner not built")


def run(*args, stdin=None, cwd=None):
    env = dict(os.environ, LANG="C", LC_ALL="C", HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    return subprocess.run([str(RUNNER), *args], input=stdin, capture_output=True, text=True, cwd=cwd or "/tmp",
                          env=env, timeout=120)


def test_list_algorithms():
    r = run("list", "algorithms")
    assert r.returncode == 0
    assert r.stdout.split() == ["bfs", "pagerank"]


def test_bfs_implementation_names():
    r = run("list", "implementations", "-a", "bfs")
    assert r.returncode == 0
    assert r.stdout.split() == sorted(list(bfs.IMPLEMENTATIONS) + ["switch", "native", "native-learned",
                                                                   "native-schedule", "native-beamer"])


def test_pagerank_implementation_names():
    r = run("list", "implementations", "-a", "pagerank")
    assert r.returncode == 0
    assert r.stdout.split() == sorted(list(pagerank.VARIANTS) + ["switch", "vertex-pull-nodiv-fused"])


def test_implementation_flags_in_help():
    r = run("list", "implementations", "-a", "bfs", "-v")
    text = r.stdout
    assert "-r NUM | --root NUM" in text and "-w NUM | --warp NUM" in text and "-m FILE | --model FILE" in text
    r = run("list", "implementations", "-a", "pagerank", "-v")
    assert "--iterations NUM" in r.stdout and "--dangling MODE" in r.stdout


def test_version_query():
    r = run("query", "algorithm-version", "-a", "bfs")
    assert r.returncode == 0 and r.stdout.strip().startswith("belewitte-amd")


def test_unknown_algorithm():
    r = run("list", "implementations", "-a", "sssp")
    assert r.returncode == 1
    assert 'No algorithm named "sssp"!' in r.stderr and "bfs" in r.stderr


def test_usage_without_arguments():
    r = run()
    assert r.returncode == 0 and "Usage:" in r.stdout


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is visible")
def test_no_device_is_an_error():
    r = run("-a", "bfs", "-k", "vertex-push", "x.graph")
    assert r.returncode == 1 and "Backend not initialised" in r.stderr
