"""What the trained network gains over the methods without it.

The shipped model (belewitte_amd/data/bfs_switch_model.json, trained on
the GPU from exhaustive per-level timings of other graphs) is evaluated on
an undirected R-MAT graph it has not seen, over all eleven variants (the
native ones and the reference's kernel families): per-level exhaustive
timings give the per-level oracle and the best fixed variant (the fallback
when no model is loaded), Beamer's direction-optimising rule is the
heuristic baseline, and the engine runs the learned policy end to end.
The switch must beat the best fixed variant clearly, stay within 10 % of the
oracle and near Beamer's rule, and -- as every variant computes the same
levels -- never change a depth."""
import numpy as np
import pytest

from belewitte_amd import bfs_native, device, graph, switch
from oracle import binding as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(gpu):
    assert switch.MODEL_PATH.exists(), "the trained switch model is missing"
    return switch.Model.load(switch.MODEL_PATH)


def test_model_matches_engine_layout(model):
    assert model.K == len(bfs_native.VARIANTS)
    prov = model.provenance
    assert prov["engine"]["sources_sha256"] and prov["graphs"] and prov["epochs"] > 0
    names = {g["graph"] for g in prov["graphs"]}
    assert "rmat-s19-ef16-u-seed11" not in names  # the evaluation graph below is held out
    assert model.w.size == switch.num_params(model.K)
    F = bfs_native.NUM_FEATURES
    assert np.all(np.isfinite(model.w)) and np.all(model.w[F:2 * F] > 0)  # positive feature scales


def test_learned_switch_beats_best_fixed(gpu, model):
    g = graph.rmat(19, 16, seed=11, undirected=True)
    dg = device.DeviceGraph(g)
    eng = bfs_native.BFSEngine(dg, edge_lists=True)
    assert eng.available == bfs_native.VARIANTS
    src = graph.random_sources(g, 4, seed=3)
    csr = O.CSR.from_graph(g)
    ref = {int(s): O.bfs_queue(csr, int(s))[0] for s in src}
    res = switch.evaluate(eng, g, src, model, repeats=2, runs=3, reference=ref)
    print(res["simulated_ms"], res["executed_ms"], flush=True)
    assert res["depth_mismatches_learned_vs_exhaustive"] == 0
    assert res["depth_mismatches_beamer_vs_exhaustive"] == 0
    assert res["depth_mismatches_vs_reference"] == 0
    sim = res["simulated_ms"]
    assert res["simulated_speedup_vs_best_fixed"] >= 1.5, sim
    assert res["simulated_regret_vs_oracle"] <= 0.10, sim
    assert sim["learned"] <= 1.10 * sim["beamer"], sim
    assert res["executed_speedup_vs_best_fixed"] >= 1.3, res["executed_ms"]
    # the learned runs themselves: exact depths from every source
    eng.set_model(model.w)
    for s in src:
        r = eng.run(int(s), "learned", "td")
        want, _ = O.bfs_queue(csr, int(s))
        assert np.array_equal(r.levels, want)
    eng.close()
    dg.free()
