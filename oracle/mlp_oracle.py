"""FP64 numpy reference of the learned LW gas-optics emulator (test
infrastructure only): forward, backward and one Adam step of the
7 -> 128 -> 128 -> 128 -> 256 softplus MLP.  Gradients are pinned by
central finite differences in tests/test_oracle_mlp.py."""
from __future__ import annotations

import numpy as np

SIZES = (7, 128, 128, 128, 256)


def softplus(z):
    return np.maximum(z, 0.0) + np.log1p(np.exp(-np.abs(z)))


def sigmoid(z):
    return 0.5 * (1.0 + np.tanh(0.5 * z))


def init_params(seed=0, sizes=SIZES):
    rng = np.random.default_rng(seed)
    p = []
    for k, n in zip(sizes[:-1], sizes[1:]):
        p.append((rng.normal(0.0, np.sqrt(2.0 / k), (k, n)), np.zeros(n)))
    return p


def forward(params, x):
    """x (rows, 7) -> y (rows, 256); also returns pre-activations / activations."""
    hs, zs = [x], []
    h = x
    for i, (W, b) in enumerate(params):
        z = h @ W + b
        zs.append(z)
        h = softplus(z) if i < len(params) - 1 else z
        hs.append(h)
    return h, hs, zs


def loss_and_grads(params, x, t):
    """Mean squared error over all rows and outputs, and its gradients."""
    y, hs, zs = forward(params, x)
    n = y.size
    d = y - t
    loss = float((d * d).sum() / n)
    g = 2.0 * d / n
    grads = [None] * len(params)
    for i in range(len(params) - 1, -1, -1):
        W, b = params[i]
        grads[i] = (hs[i].T @ g, g.sum(axis=0))
        if i > 0:
            g = (g @ W.T) * sigmoid(zs[i - 1])
    return loss, grads


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    out = []
    for i, ((W, b), (gW, gb)) in enumerate(zip(params, grads)):
        new = []
        for j, (p, g) in enumerate(((W, gW), (b, gb))):
            m[i][j] = b1 * m[i][j] + (1 - b1) * g
            v[i][j] = b2 * v[i][j] + (1 - b2) * g * g
            mh = m[i][j] / (1 - b1 ** t)
            vh = v[i][j] / (1 - b2 ** t)
            new.append(p - lr * mh / (np.sqrt(vh) + eps))
        out.append(tuple(new))
    return out
