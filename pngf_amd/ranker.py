"""Training pipeline of the learned flip ranker.

1. collect(): seeded plain DBS runs (LOWRANK, raster driver, mean |S11|^2) log, for every
   scored candidate, its ranker inputs and the exact change of |S11_f|^2 per frequency
   (pngf_dbs_params.log_x / log_y: the tuples DBS itself produces).
2. fit_norm() / train(): standardise, then train the BF16 MLP with the library's HIP
   forward / backward / Adam kernels.
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


def fit_norm(X, Y):
    xm, xs = X.mean(0), X.std(0)
    ym, ys = Y.mean(0), Y.std(0)
    xs = np.where(xs > 1e-12, xs, 1.0)
    ys = np.where(ys > 1e-30, ys, 1.0)
    return [a.astype(np.float32) for a in (xm, xs, ym, ys)]


def train(X, Y, norm, dh=256, epochs=20, batch=512, lr=1e-3, seed=0):
    xm, xs, ym, ys = norm
    Xn = ((X - xm) / xs).astype(np.float32)
    Yn = ((Y - ym) / ys).astype(np.float32)
    mlp = MLP(X.shape[1], dh, Y.shape[1], seed=seed)
    loss = mlp.train(Xn, Yn, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, loss


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
    runs = {}
    runs["steepest_exact"] = opt.dbs_run(order=ORDER_BEST)
    opt.restore()
    runs["ranked_top32"] = opt.dbs_run_ranked(mlp, norm, top_k=32, window=8, min_hit_rate=0.25)
    opt.restore()
    untrained = MLP(X.shape[1], 256, Y.shape[1], seed=7)
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
