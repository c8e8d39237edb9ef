"""The learned per-level BFS switch: data, on-GPU training, evaluation.

The reference switches BFS implementations per level through a model
looked up in SwitchImplementation (ImplementationTemplate.hpp:545-908)
and compares it with the best non-switching implementation and the
per-level optimum (benchmark-analysis/src/StepAggregate.hs:95-119,
benchmark-analysis/model-src/Evaluate.hs:229-230).  Here:

* :func:`collect` times every engine variant (the native ones and the
  reference's eight kernel families) at every level from the same saved
  state (exhaustive per-level timing) and records the level's features;
* :class:`Model` is the FP32 MLP 11 -> 32 -> 32 -> K, trained by
  :meth:`Model.train` with the hand-written GPU kernels (bw_mlp_train),
  predicting the log2 level time of each variant;
* :func:`evaluate` compares, per source, the per-level oracle (sum over
  levels of the fastest variant), the best fixed variant (the variant with
  the least total time) and the learned switch, both simulated on the
  exhaustive timings and executed end to end by the engine with the model
  evaluated on the device, beside Beamer's direction-optimising rule.
"""
from __future__ import annotations

import hashlib
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native as N
from .bfs_native import NUM_FEATURES, VARIANTS, BFSEngine, teps_edges
from .device import DeviceArray

HIDDEN = 32
TARGET_SPAN = 3.0  # log2 units: targets are capped at 8x the level's fastest time


def num_params(outputs: int = len(VARIANTS)) -> int:
    return int(N.hip_lib().bw_mlp_params(outputs))


def _layout(K: int) -> dict:
    F, H = NUM_FEATURES, HIDDEN
    o = {"mean": 0, "std": F, "W1": 2 * F}
    o["b1"] = o["W1"] + H * F
    o["W2"] = o["b1"] + H
    o["b2"] = o["W2"] + H * H
    o["W3"] = o["b2"] + H
    o["b3"] = o["W3"] + K * H
    o["end"] = o["b3"] + K
    return o


@dataclass
class Samples:
    X: np.ndarray       # float32[n][NUM_FEATURES] features
    ms: np.ndarray      # float64[n][K] level time of each variant
    source: np.ndarray  # int64[n]
    depth: np.ndarray   # int64[n]

    @classmethod
    def from_records(cls, records) -> "Samples":
        recs = list(records)
        return cls(np.array([r.features for r in recs], dtype=np.float32).reshape(-1, NUM_FEATURES),
                   np.array([r.ms for r in recs], dtype=np.float64).reshape(len(recs), -1),
                   np.array([r.source for r in recs], dtype=np.int64),
                   np.array([r.depth for r in recs], dtype=np.int64))

    @property
    def targets(self) -> np.ndarray:
        """log2 of the level time in ns (the regression target), capped at
        TARGET_SPAN above the level's fastest variant: how much slower than
        8x the best a variant runs never changes the choice, and fitting it
        would take capacity from the variants that compete."""
        y = np.log2(np.maximum(self.ms, 1e-6) * 1e6)
        if y.size:
            y = np.minimum(y, y.min(axis=1, keepdims=True) + TARGET_SPAN)
        return y.astype(np.float32)

    def concat(self, other: "Samples") -> "Samples":
        return Samples(np.concatenate([self.X, other.X]), np.concatenate([self.ms, other.ms]),
                       np.concatenate([self.source, other.source]), np.concatenate([self.depth, other.depth]))

    def save(self, path) -> None:
        np.savez(path, X=self.X, ms=self.ms, source=self.source, depth=self.depth)

    @classmethod
    def load(cls, path) -> "Samples":
        with np.load(path, allow_pickle=False) as z:
            return cls(z["X"], z["ms"], z["source"], z["depth"])


class Model:
    """Packed FP32 weights of the MLP (mean, std, W1, b1, W2, b2, W3, b3)."""

    def __init__(self, weights: np.ndarray, outputs: int = len(VARIANTS), provenance: dict | None = None):
        self.K = int(outputs)
        self.w = np.ascontiguousarray(weights, dtype=np.float32)
        self.provenance = dict(provenance or {})
        if self.w.size != num_params(self.K):
            raise ValueError(f"expected {num_params(self.K)} parameters, got {self.w.size}")

    @classmethod
    def init(cls, X: np.ndarray, outputs: int = len(VARIANTS), seed: int = 0, target_mean: float = 0.0):
        """He-initialised weights; mean/std taken from the training features."""
        K = outputs
        L = _layout(K)
        rng = np.random.default_rng(seed)
        w = np.zeros(L["end"], dtype=np.float32)
        X = np.asarray(X, dtype=np.float64)
        w[L["mean"]:L["std"]] = X.mean(axis=0) if X.size else 0.0
        sd = X.std(axis=0) if X.size else np.ones(NUM_FEATURES)
        w[L["std"]:L["W1"]] = np.where(sd > 1e-6, sd, 1.0)
        w[L["W1"]:L["b1"]] = rng.normal(0, np.sqrt(2.0 / NUM_FEATURES), HIDDEN * NUM_FEATURES)
        w[L["W2"]:L["b2"]] = rng.normal(0, np.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
        w[L["W3"]:L["b3"]] = rng.normal(0, np.sqrt(1.0 / HIDDEN), K * HIDDEN)
        w[L["b3"]:L["end"]] = target_mean
        return cls(w, K)

    # ------------------------------------------------------------ device

    def predict(self, X: np.ndarray) -> np.ndarray:
        """Predicted log2 level times, float32[n][K] (bw_mlp_forward)."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, NUM_FEATURES)
        n = X.shape[0]
        if n == 0:
            return np.zeros((0, self.K), dtype=np.float32)
        dw, dx = DeviceArray.from_host(self.w), DeviceArray.from_host(X)
        dout = DeviceArray(n * self.K, np.float32)
        N.check_hip(N.hip_lib().bw_mlp_forward(dw.ptr, self.K, dx.ptr, n, dout.ptr, None))
        out = dout.download().reshape(n, self.K)
        for a in (dw, dx, dout):
            a.free()
        return out

    def choose(self, X: np.ndarray) -> np.ndarray:
        return np.argmin(self.predict(X), axis=1)

    def gradient(self, X: np.ndarray, Y: np.ndarray):
        """(grad, loss) of the mean squared error over <= 128 samples
        (bw_mlp_grad)."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, NUM_FEATURES)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(X.shape[0], self.K)
        dw, dx, dy = DeviceArray.from_host(self.w), DeviceArray.from_host(X), DeviceArray.from_host(Y)
        dg, dl = DeviceArray(self.w.size, np.float32), DeviceArray(1, np.float32)
        N.check_hip(N.hip_lib().bw_mlp_grad(dw.ptr, self.K, dx.ptr, dy.ptr, X.shape[0], dg.ptr, dl.ptr, None))
        g, loss = dg.download(), float(dl.download()[0])
        for a in (dw, dx, dy, dg, dl):
            a.free()
        return g, loss

    def train(self, X: np.ndarray, Y: np.ndarray, epochs: int = 2000, batch: int = 64, lr: float = 3e-3,
              seed: int = 1) -> np.ndarray:
        """Trains in place on the GPU (bw_mlp_train); returns the per-epoch loss."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, NUM_FEATURES)
        Y = np.ascontiguousarray(Y, dtype=np.float32).reshape(X.shape[0], self.K)
        dw, dx, dy = DeviceArray.from_host(self.w), DeviceArray.from_host(X), DeviceArray.from_host(Y)
        dl = DeviceArray(max(epochs, 1), np.float32)
        N.check_hip(N.hip_lib().bw_mlp_train(dw.ptr, self.K, dx.ptr, dy.ptr, X.shape[0], epochs, batch,
                                             lr, seed, dl.ptr, None))
        self.w = dw.download()
        loss = dl.download()[:epochs]
        for a in (dw, dx, dy, dl):
            a.free()
        return loss

    # -------------------------------------------------------------- files

    def save(self, path) -> None:
        Path(path).write_text(json.dumps({"format": "belewitte-bfs-mlp", "outputs": self.K,
                                          "variants": list(VARIANTS[:self.K]), "hidden": HIDDEN,
                                          "features": NUM_FEATURES, "provenance": self.provenance,
                                          "weights": self.w.tolist()}))

    @classmethod
    def load(cls, path) -> "Model":
        d = json.loads(Path(path).read_text())
        if d.get("format") != "belewitte-bfs-mlp":
            raise ValueError(f"{path} is not a BFS switch model")
        return cls(np.asarray(d["weights"], dtype=np.float32), int(d["outputs"]), d.get("provenance"))


def collect(engine: BFSEngine, sources, repeats: int = 1, progress=None):
    """Exhaustive per-level timings of every variant from each source.
    Returns (Samples, {source: (levels, counts)}); `progress(i, source,
    levels)` is called after each source."""
    records, results = [], {}
    for i, s in enumerate(sources):
        levels, counts, recs = engine.exhaustive(int(s), repeats=repeats)
        records.extend(recs)
        results[int(s)] = (levels, counts)
        if progress:
            progress(i, int(s), len(recs))
    return Samples.from_records(records), results


def fit(samples: Samples, epochs: int = 100, batch: int = 64, lr: float = 3e-3, seed: int = 0):
    """A model trained on the GPU on `samples`; returns (model, epoch losses).
    Every variant must be timed at every level (no inf)."""
    if not np.all(np.isfinite(samples.ms)):
        raise ValueError("training needs every variant timed at every level")
    Y = samples.targets
    m = Model.init(samples.X, outputs=samples.ms.shape[1], seed=seed,
                   target_mean=float(Y.mean()) if Y.size else 0.0)
    loss = m.train(samples.X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
    return m, loss


def balanced(groups) -> Samples:
    """The groups concatenated, each repeated so that every group weighs
    about alike in the loss."""
    most = max(g.X.shape[0] for g in groups)
    out = None
    for g in groups:
        for _ in range(max(1, int(round(most / g.X.shape[0])))):
            out = g if out is None else out.concat(g)
    return out


def regret(samples: Samples, choices: np.ndarray) -> float:
    """Total time of the per-level choices over the per-level optimum, - 1."""
    got = samples.ms[np.arange(samples.ms.shape[0]), choices].sum()
    return float(got / samples.ms.min(axis=1).sum() - 1.0)


def cross_validate(groups, candidates=(25, 50, 100, 200, 400), folds: int = 5, seed: int = 0) -> dict:
    """Held-out-graph regret of the model trained for each epoch count in
    `candidates`: the graphs are dealt into `folds` folds (graph i to fold
    i mod folds), each fold is held out in turn, and the regret of each
    held-out graph is averaged.  Longer training fits the training graphs
    ever closer and carries over to other graphs worse."""
    folds = max(2, min(folds, len(groups)))
    table = {}
    for ep in candidates:
        r = []
        for f in range(folds):
            held = [g for i, g in enumerate(groups) if i % folds == f]
            rest = [g for i, g in enumerate(groups) if i % folds != f]
            m, _ = fit(balanced(rest), epochs=ep, seed=seed)
            r.extend(regret(h, m.choose(h.X)) for h in held)
        table[ep] = float(np.mean(r))
    return table


def simulate(samples: Samples, choices: np.ndarray) -> dict:
    """Per-source totals of the per-level oracle, each fixed variant and a
    given per-level choice, on the exhaustive timings (ms)."""
    out = {}
    for s in np.unique(samples.source):
        m = samples.source == s
        t = samples.ms[m]
        out[int(s)] = {"oracle": float(t.min(axis=1).sum()),
                       "fixed": t.sum(axis=0).tolist(),
                       "switch": float(t[np.arange(t.shape[0]), choices[m]].sum())}
    return out


def evaluate(engine: BFSEngine, g, sources, model: Model, repeats: int = 1, runs: int = 3,
             levels_per_sync: int = 8, samples_out=None, reference=None, slow_factor: float = 20.0,
             exhaustive_sources: int | None = None, progress=None) -> dict:
    """Learned switch vs best fixed vs per-level oracle vs Beamer's rule on
    `sources`, over every variant the engine can run.

    Simulated: sums of exhaustive per-level timings (device time of each
    level kernel); Beamer's choices are those of its executed runs (every
    variant reaches the same frontiers, so its levels align with the
    exhaustive ones).  Executed: whole searches run by the engine with the
    level loop on the device (hipGraph batches), min over `runs`; the
    learned policy evaluates the model on the device at every level.  A
    fixed variant slower than `slow_factor` times the best in the
    simulation is not executed (its executed time is reported as None).
    The simulation uses the first `exhaustive_sources` sources (all by
    default), the executed runs every source.  `reference` ({source:
    depths} from an independent BFS) checks the depths of the exhaustive,
    learned and Beamer runs of every source; `progress(message)` is called
    as each source is done."""
    sim_src = list(sources)[:exhaustive_sources] if exhaustive_sources else list(sources)
    say = progress or (lambda msg: None)
    samples, results = collect(engine, sim_src, repeats=repeats,
                               progress=lambda i, v, n: say(f"exhaustive source {i + 1}/{len(sim_src)} ({v}): "
                                                            f"{n} levels"))
    if samples_out:
        samples.save(samples_out)
    K = len(VARIANTS)
    pred = model.predict(samples.X)
    avail = np.array([v in engine.available for v in VARIANTS[:model.K]])
    learned_choice = np.argmin(np.where(avail[None, :], pred, np.inf), axis=1)
    sim = simulate(samples, learned_choice)
    fixed_tot = np.zeros(K)
    oracle_tot = switch_tot = 0.0
    for s in sim.values():
        fixed_tot += np.asarray(s["fixed"])
        oracle_tot += s["oracle"]
        switch_tot += s["switch"]
    best_fixed = int(np.argmin(fixed_tot))

    engine.set_model(model.w)
    exec_fixed = np.zeros(K)
    run_fixed = np.isfinite(fixed_tot) & (fixed_tot <= slow_factor * fixed_tot[best_fixed])
    exec_learned = exec_beamer = sim_beamer = 0.0
    edges = 0
    mismatches = ref_mismatches = beamer_mismatches = 0
    beamer_hist = np.zeros(K, dtype=np.int64)
    for i, s in enumerate(sources):
        exhausted = int(s) in results
        if exhausted:
            levels_ref, counts_ref = results[int(s)]
            if reference is not None and not np.array_equal(levels_ref, reference[int(s)]):
                ref_mismatches += 1
        for k in np.flatnonzero(run_fixed):
            # A variant far off the best in the simulation is run once only.
            reps = runs if fixed_tot[k] <= 2.0 * fixed_tot[best_fixed] else 1
            best = np.inf
            for _ in range(reps):
                r = engine.run(int(s), "fixed", int(k), levels_per_sync=levels_per_sync, download=False)
                best = min(best, r.total_ms)
            exec_fixed[k] += best
        best = np.inf
        for _ in range(runs):
            r = engine.run(int(s), "learned", best_fixed, levels_per_sync=levels_per_sync, download=False)
            best = min(best, r.total_ms)
        exec_learned += best
        lv = engine.levels()
        e = teps_edges(g, lv)
        edges += e
        if not exhausted:
            levels_ref, counts_ref = lv, r.counts
        else:
            if not (np.array_equal(lv, levels_ref) and np.array_equal(r.counts, counts_ref)):
                mismatches += 1
        if reference is not None and not np.array_equal(lv, reference[int(s)]):
            ref_mismatches += 1
        best = np.inf
        for _ in range(runs):
            rb = engine.run(int(s), "beamer", "td", levels_per_sync=levels_per_sync, download=False)
            best = min(best, rb.total_ms)
        exec_beamer += best
        lv = engine.levels()
        if not (np.array_equal(lv, levels_ref) and np.array_equal(rb.counts, counts_ref)):
            beamer_mismatches += 1
        if exhausted:
            m = samples.source == int(s)
            t = samples.ms[m]
            sim_beamer += float(t[np.arange(t.shape[0]), rb.choices[:t.shape[0]]].sum())
        beamer_hist += np.bincount(rb.choices, minlength=K)
        say(f"executed source {i + 1}/{len(sources)} ({int(s)}): learned {exec_learned:.3f} ms, "
            f"beamer {exec_beamer:.3f} ms so far")
    exec_fixed = np.where(run_fixed, exec_fixed, np.inf)
    exec_best = int(np.argmin(exec_fixed))
    gteps = lambda ms: edges / (ms * 1e-3) / 1e9 if ms > 0 else 0.0  # noqa: E731
    fin = lambda x: float(x) if np.isfinite(x) else None  # noqa: E731
    return {
        "sources": [int(s) for s in sources],
        "exhaustive_sources": [int(s) for s in sim_src],
        "levels": int(samples.X.shape[0]),
        "teps_edges": int(edges),
        "variants": list(engine.available),
        "simulated_ms": {"oracle": oracle_tot, "best_fixed": float(fixed_tot[best_fixed]),
                         "best_fixed_variant": VARIANTS[best_fixed], "learned": switch_tot, "beamer": sim_beamer,
                         "fixed": {v: fin(x) for v, x in zip(VARIANTS, fixed_tot)}},
        "simulated_speedup_vs_best_fixed": float(fixed_tot[best_fixed] / switch_tot) if switch_tot else None,
        "simulated_regret_vs_oracle": float(switch_tot / oracle_tot - 1.0) if oracle_tot else None,
        "simulated_beamer_regret_vs_oracle": float(sim_beamer / oracle_tot - 1.0) if oracle_tot else None,
        "executed_ms": {"best_fixed": float(exec_fixed[exec_best]), "best_fixed_variant": VARIANTS[exec_best],
                        "learned": float(exec_learned), "beamer": float(exec_beamer),
                        "fixed": {v: fin(x) for v, x in zip(VARIANTS, exec_fixed)}},
        "executed_gteps": {"learned": gteps(exec_learned), "best_fixed": gteps(exec_fixed[exec_best]),
                           "beamer": gteps(exec_beamer)},
        "executed_speedup_vs_best_fixed": float(exec_fixed[exec_best] / exec_learned) if exec_learned else None,
        "executed_speedup_vs_beamer": float(exec_beamer / exec_learned) if exec_learned else None,
        "depth_mismatches_learned_vs_exhaustive": mismatches,
        "depth_mismatches_beamer_vs_exhaustive": beamer_mismatches,
        "depth_mismatches_vs_reference": ref_mismatches if reference is not None else None,
        "learned_choice_histogram": dict(zip(VARIANTS, np.bincount(learned_choice, minlength=K).tolist())),
        "beamer_choice_histogram": dict(zip(VARIANTS, beamer_hist.tolist())),
    }


MODEL_PATH = Path(__file__).resolve().parent / "data" / "bfs_switch_model.json"


def training_graphs(small: bool = False):
    """The training inputs, across the families and scales a search meets:
    Erdos-Renyi (undirected and directed, average degree 8 to 64), 4-neighbour
    grids with 1 to 10 % of their edges rewired, and R-MAT (undirected and
    directed, edge factors 8 to 32) from 2^17 to 2^24 vertices.  The
    variants' relative speed shifts with the scale (how much of the bitmaps
    the caches hold), so every family appears at several.  small=True keeps
    three small graphs for quick runs."""
    from . import graph as G
    if small:
        yield G.erdos_renyi(1 << 16, 1 << 19, seed=3, undirected=True)
        yield G.grid2d(256, 256, rewire=0.01, seed=4)
        yield G.rmat(16, 16, seed=2)
        return
    yield G.erdos_renyi(1 << 24, 134217728, seed=3, undirected=True)
    yield G.grid2d(4096, 4096, rewire=0.01, seed=4)
    yield G.rmat(22, 16, seed=2)
    yield G.rmat(23, 16, seed=25, undirected=True)
    yield G.rmat(24, 16, seed=28)
    yield G.rmat(21, 16, seed=21, undirected=True)
    yield G.erdos_renyi(1 << 21, 1 << 26, seed=23, undirected=True)
    yield G.grid2d(2048, 2048, rewire=0.05, seed=24)
    yield G.erdos_renyi(1 << 22, 1 << 25, seed=26, undirected=False)
    yield G.rmat(20, 8, seed=22)
    yield G.rmat(18, 32, seed=27, undirected=True)
    yield G.rmat(17, 16, seed=29, undirected=True)
    yield G.rmat(18, 16, seed=30)
    yield G.erdos_renyi(1 << 18, 1 << 21, seed=31, undirected=True)
    yield G.erdos_renyi(1 << 19, 1 << 22, seed=32, undirected=False)
    yield G.grid2d(1024, 1024, rewire=0.01, seed=33)
    yield G.grid2d(512, 512, rewire=0.1, seed=34)
    yield G.rmat(20, 16, seed=35, undirected=True)
    yield G.erdos_renyi(1 << 20, 1 << 24, seed=36, undirected=True)


ENGINE_SOURCES = ("csrc/hip/bfs_native.hip", "csrc/hip/mlp.hip", "../include/belewitte_hip.h")


def engine_fingerprint() -> dict:
    """What identifies the engine a model was trained on: the library's
    build version and a SHA-256 over the engine's sources."""
    h = hashlib.sha256()
    here = Path(__file__).resolve().parent
    for rel in ENGINE_SOURCES:
        h.update((here / rel).read_bytes())
    return {"library": N.hip_lib().bw_hip_version().decode(), "sources_sha256": h.hexdigest(),
            "sources": [str(Path(r).name) for r in ENGINE_SOURCES]}


def main(argv=None) -> int:
    import argparse
    import sys
    import time

    from . import device
    from . import graph as G
    ap = argparse.ArgumentParser(description="collect exhaustive per-level BFS timings and train the switch")
    ap.add_argument("--out", default=str(MODEL_PATH))
    ap.add_argument("--samples-out", default=None, help="save the exhaustive timings (.npz, graphs in .npz.json)")
    ap.add_argument("--samples-in", default=None, help="train on saved timings instead of collecting them")
    ap.add_argument("--sources", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=0, help="training epochs (0: chosen by cross-validation)")
    ap.add_argument("--small", action="store_true")
    args = ap.parse_args(argv)
    device.require_gpu()
    device.set_device(0)
    groups, trained_on = [], []
    if args.samples_in:
        saved = Samples.load(args.samples_in)
        trained_on = json.loads(Path(args.samples_in + ".json").read_text())
        for gi in range(len(trained_on)):
            m = (saved.source >> 32) == gi
            groups.append(Samples(saved.X[m], saved.ms[m], saved.source[m], saved.depth[m]))
    for g in ([] if args.samples_in else training_graphs(args.small)):
        t0 = time.perf_counter()
        print(f"[switch] {g.name} generated", file=sys.stderr, flush=True)
        dg = device.DeviceGraph(g)
        eng = BFSEngine(dg, edge_lists=True)
        src = G.random_sources(g, args.sources, seed=5)
        s, _ = collect(eng, src, progress=lambda i, v, n: print(
            f"[switch] {g.name}: source {i + 1}/{len(src)} ({v}): {n} levels", file=sys.stderr, flush=True))
        print(f"[switch] {g.name}: {s.X.shape[0]} levels from {len(src)} sources in "
              f"{time.perf_counter() - t0:.1f}s; best-fixed totals {s.ms.sum(axis=0).round(3).tolist()} ms, "
              f"oracle {s.ms.min(axis=1).sum():.3f} ms", file=sys.stderr, flush=True)
        s.source = s.source + (np.int64(len(groups)) << 32)  # sources unique across graphs
        groups.append(s)
        trained_on.append({"graph": g.name, "vertices": int(g.vertex_count), "edges": int(g.edge_count),
                           "sources": len(src), "source_seed": 5, "levels": int(s.X.shape[0])})
        eng.close()
        dg.free()
    samples = groups[0]
    for s in groups[1:]:
        samples = samples.concat(s)
    if args.samples_out:
        samples.save(args.samples_out)
        Path(args.samples_out + ".json").write_text(json.dumps(trained_on))
    epochs = args.epochs
    cv = None
    if epochs <= 0:
        cv = cross_validate(groups)
        epochs = min(cv, key=cv.get)
        print(f"[switch] held-out-graph regret by epochs: "
              f"{ {k: round(v, 4) for k, v in cv.items()} }; training {epochs} epochs", file=sys.stderr, flush=True)
    model, loss = fit(balanced(groups), epochs=epochs)
    choice = model.choose(samples.X)
    sim = simulate(samples, choice)
    fixed = np.sum([v["fixed"] for v in sim.values()], axis=0)
    learned = sum(v["switch"] for v in sim.values())
    oracle = sum(v["oracle"] for v in sim.values())
    print(f"[switch] trained: loss {loss[0]:.4f} -> {loss[-1]:.4f}; training-set totals: oracle {oracle:.3f} ms, "
          f"learned {learned:.3f} ms, best fixed {fixed.min():.3f} ms ({VARIANTS[int(np.argmin(fixed))]})",
          file=sys.stderr, flush=True)
    import datetime
    props = device.device_props(0)
    model.provenance = {
        "engine": engine_fingerprint(),
        "device": f"{props['arch']}, {props['compute_units']} CUs" + (f" ({props['name']})" if props["name"] else ""),
        "trained": datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"),
        "graphs": trained_on,
        "epochs": int(epochs), "seed": 0,
        "cross_validation_regret": {str(k): v for k, v in cv.items()} if cv else None,
        "training_set_ms": {"oracle": float(oracle), "learned": float(learned),
                            "fixed": dict(zip(VARIANTS, np.asarray(fixed, dtype=float).tolist()))},
        "command": "python -m belewitte_amd.switch" + (" --small" if args.small else "")
                   + f" --sources {args.sources}" + (f" --epochs {args.epochs}" if args.epochs else "")
                   + (" (trained on saved timings)" if args.samples_in else ""),
    }
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    model.save(args.out)
    print(args.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
