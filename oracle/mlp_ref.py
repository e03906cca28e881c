"""CPU reference of the learned flip ranker -- TEST INFRASTRUCTURE ONLY.

Forward and backward of the din -> dh -> dh -> dout GELU network in float64, with
every matrix-product operand rounded to BF16 (round to nearest even) exactly where the
HIP kernels round them, so that the GPU gradients can be compared element-wise.
``bf16=False`` gives the smooth float64 network used to pin the backward formulas by
central finite differences.
"""
import numpy as np


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


C_GELU = np.sqrt(2.0 / np.pi)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    u = C_GELU * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C_GELU * (1.0 + 3 * 0.044715 * x * x)


def split(w, din, dh, dout):
    sizes = [din * dh, dh, dh * dh, dh, dh * dout, dout]
    shapes = [(din, dh), (dh,), (dh, dh), (dh,), (dh, dout), (dout,)]
    out, off = [], 0
    for n, s in zip(sizes, shapes):
        out.append(np.asarray(w[off:off + n], dtype=np.float64).reshape(s))
        off += n
    return out


def forward(w, X, din, dh, dout, rnd=True):
    r = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(w, din, dh, dout)
    Z1 = r(X) @ r(W1) + b1
    H1 = gelu(Z1)
    Z2 = r(H1) @ r(W2) + b2
    H2 = gelu(Z2)
    O = r(H2) @ r(W3) + b3
    return O, (X, Z1, H1, Z2, H2)


def loss_and_grad(w, X, Y, din, dh, dout, rnd=True):
    """Mean squared error and its gradient, in the parameter layout of the HIP MLP."""
    r = bf16 if rnd else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(w, din, dh, dout)
    O, (X, Z1, H1, Z2, H2) = forward(w, X, din, dh, dout, rnd)
    n = O.size
    D = O - Y
    loss = float(np.sum(D * D) / n)
    dO = 2.0 * D / n
    if rnd:
        # the kernels store dO, dH, dZ in FP32 before they enter the next product
        dO = dO.astype(np.float32).astype(np.float64)
    gW3 = r(H2).T @ r(dO)
    gb3 = dO.sum(0)
    dH2 = r(dO) @ r(W3).T
    dZ2 = dH2 * gelu_grad(Z2)
    gW2 = r(H1).T @ r(dZ2)
    gb2 = dZ2.sum(0)
    dH1 = r(dZ2) @ r(W2).T
    dZ1 = dH1 * gelu_grad(Z1)
    gW1 = r(X).T @ r(dZ1)
    gb1 = dZ1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g
