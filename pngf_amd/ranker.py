"""Training pipeline of the learned flip ranker.

1. collect(): seeded plain DBS runs (LOWRANK, raster driver, mean |S11|^2) log, for every
   scored candidate, its ranker inputs and the exact change of |S11_f|^2 per frequency
   (pngf_dbs_params.log_x / log_y: the tuples DBS itself produces).
2. fit_norm() / train(): standardise the inputs, standardise the labels within each state
   (the ranking inside a state is what matters), then train the BF16 MLP with the library's
   HIP forward / backward / Adam kernels.
3. Optimizer.dbs_run_ranked(): the network ranks all candidates, exact FP64 Woodbury scores
   them in that order K at a time until a chunk holds an improving flip (its best is
   accepted), full sweeps take over when the hit rate of the first chunks drops, and a step
   that exhausts the ranking is the exact check of convergence.

``python -m pngf_amd.ranker`` runs the 7x7-tile, 9-frequency workload end to end and
prints one JSON line (tuples, training loss, and exact evaluations / objective of ranked
versus unranked DBS at 21x21 tiles).
"""
import json
import sys
import time

import numpy as np

from . import LAYOUT_DISJOINT, METHOD_LOWRANK, MLP, ORDER_BEST, Optimizer
from .problem import freq_list, initial_pattern


def make_optimizer(T, freqs, seed, layout=LAYOUT_DISJOINT, synth_soft=0.0, max_deferred=0):
    opt = Optimizer(layout, T, num_freq=len(freqs), method=METHOD_LOWRANK, synth_soft=synth_soft,
                    max_deferred=max_deferred)
    for f, hz in enumerate(freqs):
        opt.set_frequency(f, hz)
        opt.synth_g(f)
    opt.set_tile_map(initial_pattern(T, seed))
    mid = (T - 1) // 2
    opt.set_source_port(mid, mid, 1, 2)
    opt.init_state()
    return opt


def collect(T, freqs, seeds, synth_soft=0.0, log_cap=100000):
    """(X, Y) tuples from one plain DBS run per seed."""
    X, Y = [], []
    for s in seeds:
        opt = make_optimizer(T, freqs, s, synth_soft=synth_soft)
        r = opt.dbs_run(log_cap=log_cap)
        X.append(r["log_x"])
        Y.append(r["log_y"])
        opt.close()
    return np.concatenate(X), np.concatenate(Y)


def state_ids(X, nf):
    """Rows scored against the same optimizer state share the current-Zin inputs (the last two
    per-frequency inputs); returns one integer id per row."""
    key = np.concatenate([X[:, 51 + 6::8], X[:, 51 + 7::8]], axis=1).astype(np.float64)
    _, ids = np.unique(key, axis=0, return_inverse=True)
    return ids.ravel()


def standardize_labels(X, Y):
    """Labels centred and scaled within each state and frequency: the ranking of the
    candidates of one state is what the driver uses, and the spread of the exact changes
    varies by orders of magnitude between states."""
    ids = state_ids(X, Y.shape[1])
    Yn = np.zeros_like(Y, dtype=np.float32)
    for s in np.unique(ids):
        m = ids == s
        if m.sum() < 2:
            continue
        y = Y[m].astype(np.float64)
        Yn[m] = (y - y.mean(0)) / (y.std(0) + 1e-30)
    return Yn


def fit_norm(X, Y):
    """Input standardisation; outputs are within-state standardised scores (mean 0, std 1)."""
    xm, xs = X.mean(0), X.std(0)
    xs = np.where(xs > 1e-12, xs, 1.0)
    nf = Y.shape[1]
    return [a.astype(np.float32) for a in (xm, xs, np.zeros(nf), np.ones(nf))]


def train(X, Y, norm, dh=256, epochs=20, batch=512, lr=1e-3, seed=0):
    xm, xs, _, _ = norm
    Xn = ((X - xm) / xs).astype(np.float32)
    Yn = standardize_labels(X, Y)
    mlp = MLP(X.shape[1], dh, Y.shape[1], seed=seed)
    loss = mlp.train(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, loss


def ranking_quality(opt, mlp, norm, top_k, max_steps=60):
    """Along the steepest-descent trajectory from opt's current state (each step accepts the
    exact best flip), compare the network's order with the exact changes of the objective:
    mean Spearman correlation, improving flips among the network's top_k (and the number a
    random order would place there), and how often the exact best is in the top_k.
    Leaves opt at the trajectory's end."""
    from scipy.stats import spearmanr
    T = opt.T
    cands = list(range(T * T))
    xm, xs, ym, ys = norm
    rho, imp, rnd, best_in = [], 0, 0.0, 0
    for _ in range(max_steps):
        F0 = opt.state()["F"]
        Fx, _ = opt.eval_batch([[t] for t in cands])
        d = Fx - F0
        nimp = int((d < 0).sum())
        if nimp == 0:
            break
        s = (mlp.predict(((opt.features(cands) - xm) / xs).astype(np.float32)) * ys + ym).mean(1)
        order = np.argsort(s, kind="stable")
        rho.append(spearmanr(s, d).correlation)
        imp += int((d[order[:top_k]] < 0).sum())
        rnd += nimp * top_k / len(cands)
        best = int(np.argmin(d))
        best_in += int(best in order[:top_k])
        opt.accept([cands[best]])
    return dict(steps=len(rho), rho=float(np.mean(rho)) if rho else 0.0, improving_in_top_k=imp,
                random_expectation=rnd, best_in_top_k=best_in)


def hit_rate(r, ncand=None):
    """Fraction of ranked steps whose first top-K chunk held an improving flip."""
    return r["ranked_hits"] / r["ranked_steps"] if r["ranked_steps"] else 0.0


def main():
    t0 = time.time()
    freqs = freq_list(9, 26e9, 34e9)
    nseed = int(sys.argv[1]) if len(sys.argv) > 1 else 64
    X, Y = collect(7, freqs, range(nseed))
    norm = fit_norm(X, Y)
    mlp, loss = train(X, Y, norm, epochs=30)
    out = {"tuples": int(len(X)), "train_loss": loss, "collect_train_s": time.time() - t0}
    # deploy on the 21x21-tile, 9-frequency problem
    opt = make_optimizer(21, freqs, 0)
    opt.checkpoint()
    untrained = MLP(X.shape[1], 256, Y.shape[1], seed=7)
    out["ranking_quality_top32"] = ranking_quality(opt, mlp, norm, 32)
    opt.restore()
    out["ranking_quality_top32_untrained"] = ranking_quality(opt, untrained, norm, 32)
    opt.restore()
    runs = {}
    runs["steepest_exact"] = opt.dbs_run(order=ORDER_BEST)
    opt.restore()
    runs["ranked_top32"] = opt.dbs_run_ranked(mlp, norm, top_k=32, window=8, min_hit_rate=0.25)
    opt.restore()
    runs["ranked_top32_untrained"] = opt.dbs_run_ranked(untrained, norm, top_k=32, window=8, min_hit_rate=0.25)
    opt.restore()
    runs["first_improvement"] = opt.dbs_run()
    for name, r in runs.items():
        out[name] = dict(exact_evals=r["flip_count"], accepts=r["accepts"], F_final=r["F_final"],
                         seconds=r["seconds"], full_sweeps=r["full_sweeps"])
        if name.startswith("ranked"):
            out[name]["hit_rate"] = hit_rate(r)
    out["ranked_F_vs_steepest"] = runs["ranked_top32"]["F_final"] / runs["steepest_exact"]["F_final"]
    out["ranked_evals_vs_steepest"] = runs["ranked_top32"]["flip_count"] / runs["steepest_exact"]["flip_count"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
