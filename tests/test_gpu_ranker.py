"""The learned flip ranker on the GPU: network kernels against the CPU reference
(oracle/mlp_ref.py), the DBS training log against the oracle, and what the trained
network gains in ranked DBS."""
import numpy as np
import pytest

import oracle as O
import pngf_amd as P
from oracle import mlp_ref as M
from pngf_amd import ranker
from pngf_amd.problem import freq_list, initial_pattern

pytestmark = pytest.mark.gpu

DX, DY, DZ = 0.01 / 60, 0.01 / 60, 0.01 / 180


def test_mlp_forward_and_gradient_match_reference():
    din, dh, dout, n = 67, 32, 2, 37
    mlp = P.MLP(din, dh, dout, seed=3)
    rng = np.random.default_rng(1)
    w = (rng.standard_normal(mlp.nparam) * 0.2).astype(np.float32)
    mlp.set_params(w)
    X = rng.standard_normal((n, din)).astype(np.float32)
    Y = rng.standard_normal((n, dout)).astype(np.float32)
    O_ref, _ = M.forward(w, X, din, dh, dout)
    O_gpu = mlp.predict(X)
    assert np.linalg.norm(O_gpu - O_ref) / np.linalg.norm(O_ref) < 1e-3
    loss_ref, g_ref = M.loss_and_grad(w, X, Y, din, dh, dout)
    g_gpu, loss_gpu = mlp.grad(X, Y)
    assert abs(loss_gpu - loss_ref) < 1e-3 * loss_ref
    for a, b in zip(mlp.split(g_gpu), M.split(g_ref, din, dh, dout)):
        assert np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12) < 2e-3


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((4096, 20)).astype(np.float32)
    a = rng.standard_normal((20, 3))
    Y = np.tanh(X @ a).astype(np.float32)
    mlp = P.MLP(20, 64, 3, seed=0)
    _, l0 = mlp.grad(X[:512], Y[:512])
    loss = mlp.train(X, Y, epochs=30, batch=256, lr=3e-3)
    assert loss < 0.1 * l0


def test_dbs_training_log_matches_oracle():
    T, freqs, soft = 7, freq_list(2, 29e9, 31e9), 0.4
    opt = ranker.make_optimizer(T, freqs, 0, synth_soft=soft)
    din, dout = opt.ranker_dims()
    assert (din, dout) == (51 + 8 * len(freqs), len(freqs))
    # the first scored batch is the raster cycle against the initial state
    cands = list(range(T * T))
    X0 = opt.features(cands)
    tm = initial_pattern(T, 0)
    for t in cands:
        assert X0[t, 24] == tm[t]  # centre of the 7x7 neighbourhood is the tile itself
    r = opt.dbs_run(log_cap=100000)
    assert len(r["log_x"]) == r["flip_count"] + r["speculative_evals"]
    assert np.allclose(r["log_x"][:5], X0[:5])
    ors = []
    for hz in freqs:
        G = P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, hz, soft=soft)
        o = O.Optimizer(P.LAYOUT_DISJOINT, T, G)
        assert o.init_state(tm, (3, 3, 1, 2)) == 0
        ors.append(o)
    V = np.array([o.scalars()["Vinc"] for o in ors])
    e0 = np.array([o.scalars()["Eobj"] for o in ors])
    _, s0, _ = O.objective(O.OBJ_S11, e0, V, None, freqs, DX, DY, DZ)
    for t in range(5):
        e = np.array([o.compute_updated_Eobj([t]) for o in ors])
        _, s, _ = O.objective(O.OBJ_S11, e, V, None, freqs, DX, DY, DZ)
        want = np.abs(s) ** 2 - np.abs(s0) ** 2
        assert np.allclose(r["log_y"][t], want, rtol=1e-5, atol=1e-9)


def test_trained_ranker_gains():
    """What the trained network gains, on three starts it was not trained on (7x7 tiles, 3
    frequencies; training tuples logged by plain DBS from 16 other starts):
    * ranking: along each start's steepest-descent trajectory the network's order correlates
      with the exact objective changes (mean Spearman rho) where an untrained network's does
      not, it places more improving flips in its top 6 than a random order would, and it
      holds the exact best flip in its top 6 more often than the untrained network;
    * ranked DBS against the method without it (exact full sweeps, best-improvement DBS):
      fewer exact FP64 evaluations for a final objective within 2 % (mean over the starts).
    Every accepted flip is an exact evaluation and every run ends in an exact local minimum."""
    T, freqs = 7, freq_list(3, 28e9, 32e9)
    X, Y = ranker.collect(T, freqs, range(1, 17))
    assert len(X) > 5000
    norm = ranker.fit_norm(X, Y)
    mlp, loss = ranker.train(X, Y, norm, dh=64, epochs=40, batch=256, lr=2e-3)
    untrained = P.MLP(X.shape[1], 64, Y.shape[1], seed=5)
    qg, qb, ev_full, ev_good, F_full, F_good = [], [], 0, 0, [], []
    for seed in (11, 12, 13):
        opt = ranker.make_optimizer(T, freqs, seed)
        opt.checkpoint()
        qg.append(ranker.ranking_quality(opt, mlp, norm, 6))
        opt.restore()
        qb.append(ranker.ranking_quality(opt, untrained, norm, 6))
        opt.restore()
        full = opt.dbs_run(order=P.ORDER_BEST)
        opt.restore()
        good = opt.dbs_run_ranked(mlp, norm, top_k=6, window=4, min_hit_rate=0.25)
        assert full["converged"] and good["converged"]
        assert np.all(np.diff(good["acc_F"]) < 0)
        ev_full += full["flip_count"]
        ev_good += good["flip_count"]
        F_full.append(full["F_final"])
        F_good.append(good["F_final"])
        opt.close()
    rho_g = np.mean([q["rho"] for q in qg])
    rho_b = np.mean([q["rho"] for q in qb])
    assert rho_g > 0.15 and rho_g > rho_b + 0.15
    assert sum(q["improving_in_top_k"] for q in qg) > sum(q["random_expectation"] for q in qg)
    assert sum(q["best_in_top_k"] for q in qg) > sum(q["best_in_top_k"] for q in qb)
    assert ev_good < 0.85 * ev_full
    assert np.mean(F_good) <= 1.02 * np.mean(F_full)


def test_best_order_reaches_a_local_minimum():
    """Steepest-descent DBS: every accept is the best exact score of its sweep and the
    run stops when a full sweep has no improving flip (checked against the oracle)."""
    T, freqs, soft = 7, freq_list(2, 29e9, 31e9), 0.4
    opt = ranker.make_optimizer(T, freqs, 4, synth_soft=soft)
    r = opt.dbs_run(order=P.ORDER_BEST)
    assert r["converged"] and r["accepts"] > 0 and np.all(np.diff(r["acc_F"]) < 0)
    fm = opt.tile_map()
    ors = []
    for hz in freqs:
        G = P.synth_g_host(P.LAYOUT_DISJOINT, T, 1, hz, soft=soft)
        o = O.Optimizer(P.LAYOUT_DISJOINT, T, G)
        assert o.init_state(fm, (3, 3, 1, 2)) == 0
        ors.append(o)
    V = np.array([o.scalars()["Vinc"] for o in ors])
    e0 = np.array([o.scalars()["Eobj"] for o in ors])
    F0, _, _ = O.objective(O.OBJ_S11, e0, V, None, freqs, DX, DY, DZ)
    assert abs(F0 - r["F_final"]) < 1e-8 * F0
    for t in range(T * T):
        e = np.array([o.compute_updated_Eobj([t]) for o in ors])
        F, _, _ = O.objective(O.OBJ_S11, e, V, None, freqs, DX, DY, DZ)
        assert F >= F0 * (1 - 1e-12)
