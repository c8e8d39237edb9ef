"""Pins oracle_pagerank_binned_fixed: the binned iteration's fixed-point
arithmetic restated on the CPU, against which the GPU test checks the
device's DROP-mode ranks bit for bit.

It is pinned by:
  * an exact brute force in Python (integer fixed-point sums, rational
    arithmetic for the fused multiply-add and the division, each rounded to
    FP32 by hand), bit for bit on small graphs, repeated edges included.
    The same brute force with one planted mistake (a multiply and an add
    rounded separately instead of one fused rounding) disagrees with it;
  * the FP64 recurrence of the same PageRank (oracle_pagerank_f64): within
    1e-6 relative on every vertex, the FP32 rounding of the sums and
    ranks (the fixed-point truncation adds at most n * 2^-S to a sum of n
    terms);
  * the reference's sequential FP32 restatement (oracle_pagerank_f32,
    vertex-pull or vertex-pull-nodiv): within 1e-6 relative beyond that
    restatement's own error against FP64.
"""
from fractions import Fraction

import numpy as np
import pytest

from belewitte_amd import graph, pagerank
from graphfiles import write_v1
from oracle import binding as O

DAMP = np.float32(0.85)


def _f32(x: Fraction) -> np.float32:
    """x >= 0 rounded to the nearest FP32 value, ties to even (normal range)."""
    if x == 0:
        return np.float32(0.0)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    k = 23 - e  # x * 2^k in [2^23, 2^24)
    scaled = x * Fraction(2) ** k
    m = scaled.numerator // scaled.denominator
    rem = scaled - m
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and m % 2 == 1):
        m += 1
    return np.float32(float(Fraction(m) / Fraction(2) ** k))


def _brute(csr: O.CSR, rows_in, deg, iterations: int, quirk: bool, fused: bool = True):
    V, E = csr.V, int(sum(len(r) for r in rows_in))
    lg = 0
    if quirk:
        M = Fraction(1) + Fraction(E, V)
        while Fraction(2) ** lg < M:
            lg += 1
    S = 62 - lg
    init = _f32(Fraction(1) / Fraction(V))
    base = _f32((Fraction(1) - Fraction(float(DAMP))) / Fraction(V))
    assert base == (np.float32(1.0) - DAMP) / np.float32(V)
    c = [init if (quirk or deg[v] == 0) else _f32(Fraction(float(init)) / deg[v]) for v in range(V)]
    ranks = [init] * V
    for _ in range(iterations):
        nc = []
        for v in range(V):
            acc = sum((Fraction(float(c[u])) * 2 ** S).__floor__() for u in rows_in[v])
            s = _f32(Fraction(acc) / Fraction(2) ** S)
            if fused:
                nr = _f32(Fraction(float(DAMP)) * Fraction(float(s)) + Fraction(float(base)))
            else:  # the planted mistake: two roundings
                nr = _f32(Fraction(float(_f32(Fraction(float(DAMP)) * Fraction(float(s))))) + Fraction(float(base)))
            ranks[v] = nr
            nc.append(_f32(Fraction(float(nr)) / deg[v]) if deg[v] else np.float32(0.0))
        c = nc
    return np.array(ranks, dtype=np.float32)


def _small_graphs(tmp_path):
    rng = np.random.default_rng(5)
    out = [graph.rmat(5, 4, seed=3), graph.star(9), graph.chain(7)]
    src = rng.integers(0, 12, 90)
    dst = np.where(rng.random(90) < 0.5, 0, rng.integers(0, 12, 90))
    write_v1(tmp_path / "dup.graph", 12, src, dst)
    out.append(graph.load(tmp_path / "dup.graph"))
    return out


def _rows_in(csr: O.CSR):
    roff, rcol = csr.roff, csr.rcol
    return [list(rcol[roff[v]:roff[v + 1]]) for v in range(csr.V)]


@pytest.mark.parametrize("quirk", [False, True])
def test_binned_fixed_matches_exact_brute_force(tmp_path, quirk):
    planted_caught = False
    for g in _small_graphs(tmp_path):
        csr = O.CSR.from_graph(g)
        rows = _rows_in(csr)
        deg = np.bincount(np.concatenate([np.asarray(r, dtype=np.int64) for r in rows] + [np.zeros(0, np.int64)]),
                          minlength=csr.V)
        for iters in (1, 3):
            want = _brute(csr, rows, deg, iters, quirk)
            got = O.pagerank_binned_fixed(csr, iters, nodiv_quirk=quirk)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (g.name, iters, got, want)
            bad = _brute(csr, rows, deg, iters, quirk, fused=False)
            planted_caught |= not np.array_equal(got.view(np.uint32), bad.view(np.uint32))
    assert planted_caught, "the unfused brute force never differed: the check would miss that mistake"


@pytest.mark.parametrize("quirk,iters", [(False, 1), (False, 20), (True, 1), (True, 12)])
def test_binned_fixed_within_fp32_of_fp64(quirk, iters):
    for g in (graph.rmat(12, 16, seed=31), graph.erdos_renyi(3000, 20000, seed=4, undirected=False),
              graph.star(300)):
        csr = O.CSR.from_graph(g)
        got = O.pagerank_binned_fixed(csr, iters, nodiv_quirk=quirk).astype(np.float64)
        x = O.pagerank_f64(csr, iters, "drop", nodiv_quirk=quirk)
        rel = float(np.max(np.abs(got - x) / x))
        assert rel <= 1e-6, (g.name, quirk, iters, rel)
        # The reference's sequential FP32 sums: vertex-pull-nodiv carries
        # the undivided first iteration, vertex-pull does not.
        ref, _ = O.pagerank_f32(csr, pagerank.variant_id("vertex-pull-nodiv" if quirk else "vertex-pull"), iters,
                                "drop")
        ref_err = np.abs(ref.astype(np.float64) - x) / x
        # Within 1e-6 of the reference's ranks beyond their own error, and
        # never less accurate than them by more than 1e-6.
        assert float(np.max(np.abs(got - ref.astype(np.float64)) / x - ref_err)) <= 1e-6, (g.name, quirk, iters)
        assert rel <= float(np.max(ref_err)) + 1e-6
