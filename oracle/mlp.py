"""Oracle MLP (test infrastructure): numpy restatement of the library's BF16
network used by the learned warm start (no counterpart in the reference).

Float64 arithmetic with BF16 round-to-nearest-even applied exactly where the
device rounds: layer inputs and weights before each product, GELU outputs
between layers, and dZ before the two backward products. With bf16=False
nothing is rounded, which the tests use to check the backward formulas
against central differences of the loss.
"""

import numpy as np
from scipy.special import erf


def to_bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    x32 = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unflatten(params, dims):
    """[(W_l (dims[l+1], dims[l]), b_l)] from the library's flat layout."""
    out, o = [], 0
    for l in range(len(dims) - 1):
        n, k = dims[l + 1], dims[l]
        W = np.asarray(params[o:o + n * k], dtype=np.float64).reshape(n, k)
        o += n * k
        b = np.asarray(params[o:o + n], dtype=np.float64)
        o += n
        out.append((W, b))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def forward(layers, X, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    A = r(X)
    acts, zs = [], []
    for l, (W, b) in enumerate(layers):
        acts.append(A)
        Z = A @ r(W).T + b
        zs.append(Z)
        if l + 1 < len(layers):
            A = r(gelu(Z))
    return zs[-1], acts, zs


def loss_and_grad(layers, X, T, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    Y, acts, zs = forward(layers, X, bf16)
    B, dL = T.shape
    loss = float(((Y - T) ** 2).sum() / (B * dL))
    dZ = 2.0 * (Y - T) / (B * dL)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dW = r(dZ).T @ acts[l]
        db = dZ.sum(axis=0)
        grads[l] = (dW, db)
        if l > 0:
            dA = r(dZ) @ r(W)
            dZ = dA * gelu_grad(zs[l - 1])
    return loss, grads


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update (step counts from 1); returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
