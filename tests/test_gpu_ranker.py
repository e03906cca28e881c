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
    """Ranked DBS with the trained network against (a) exact full sweeps every step
    (best-improvement DBS: the method without the network) and (b) the same ranked driver
    with an untrained network.  Every accepted flip is exact and each run ends in an exact
    local minimum; the trained ranking needs fewer exact evaluations than full sweeps for an
    objective within 1 %, and finds an improving flip in its first chunk more often than an
    untrained one."""
    T, freqs = 7, freq_list(3, 28e9, 32e9)
    X, Y = ranker.collect(T, freqs, range(1, 9))
    assert len(X) > 1000
    norm = ranker.fit_norm(X, Y)
    mlp, loss = ranker.train(X, Y, norm, dh=64, epochs=40, batch=256, lr=2e-3)
    untrained = P.MLP(X.shape[1], 64, Y.shape[1], seed=5)
    opt = ranker.make_optimizer(T, freqs, 11)
    opt.checkpoint()
    full = opt.dbs_run(order=P.ORDER_BEST)
    opt.restore()
    good = opt.dbs_run_ranked(mlp, norm, top_k=6, window=4, min_hit_rate=0.25)
    opt.restore()
    bad = opt.dbs_run_ranked(untrained, norm, top_k=6, window=4, min_hit_rate=0.25)
    assert good["converged"] and full["converged"] and bad["converged"]
    assert good["flip_count"] < full["flip_count"]
    assert good["F_final"] <= 1.01 * full["F_final"]
    assert ranker.hit_rate(good) > ranker.hit_rate(bad)


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
