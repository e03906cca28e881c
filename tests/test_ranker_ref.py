"""Pins the CPU reference of the ranker network (oracle/mlp_ref.py) on the CPU."""
import numpy as np

from oracle import mlp_ref as M


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0, 0.0], dtype=np.float32)
    # BF16 spacing at 1 is 2^-7: 1 + 2^-8 is a tie and goes to the even 1.0;
    # 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    assert M.bf16(x).tolist() == [1.0, 1.0, 1.0078125, -3.0, 0.0]


def test_gelu_grad_finite_difference():
    x = np.linspace(-4, 4, 101)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(x))) < 1e-8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    din, dh, dout, n = 7, 6, 3, 5
    npar = din * dh + dh + dh * dh + dh + dh * dout + dout
    w = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((n, din))
    Y = rng.standard_normal((n, dout))
    _, g = M.loss_and_grad(w, X, Y, din, dh, dout, rnd=False)
    h = 1e-6
    fd = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (M.loss_and_grad(w + e, X, Y, din, dh, dout, rnd=False)[0]
                 - M.loss_and_grad(w - e, X, Y, din, dh, dout, rnd=False)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
