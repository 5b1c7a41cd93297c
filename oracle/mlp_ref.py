"""CPU reference of the cost-model MLP (TEST INFRASTRUCTURE ONLY).

The network the library trains and runs in HIP kernels: inputs of 5 job
features, two hidden layers of 64 ReLU units, one output per candidate
configuration (a predicted log time per prime), fp32:

    h1 = relu(x W1 + b1);  h2 = relu(h1 W2 + b2);  y = h2 W3 + b3
    loss = mean over samples and outputs of mask * (y - t)^2 / 2,
           divided by the number of unmasked entries

This module computes the forward pass, the loss and the exact gradients in
float64 numpy, so that tests can check the HIP forward/backward kernels
against it.  It is pinned by finite differences (tests/test_mlp.py).  Nothing
in the product imports it.
"""
import numpy as np


def init(n_in, n_hidden, n_out, seed):
    """Deterministic He-uniform initialisation (the library uses the same
    formula with its own generator; tests pass weights explicitly)."""
    rng = np.random.default_rng(seed)
    def lin(a, b):
        lim = np.sqrt(6.0 / a)
        return rng.uniform(-lim, lim, size=(a, b)).astype(np.float32), np.zeros(b, np.float32)
    W1, b1 = lin(n_in, n_hidden)
    W2, b2 = lin(n_hidden, n_hidden)
    W3, b3 = lin(n_hidden, n_out)
    return [W1, b1, W2, b2, W3, b3]


def forward(params, X):
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    X = np.asarray(X, np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ W3 + b3
    return y, (X, z1, h1, z2, h2)


def loss_and_grads(params, X, T, M=None):
    """Masked mean squared error / 2 and its gradients w.r.t. every parameter."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    y, (X, z1, h1, z2, h2) = forward(params, X)
    T = np.asarray(T, np.float64)
    M = np.ones_like(T) if M is None else np.asarray(M, np.float64)
    cnt = max(M.sum(), 1.0)
    r = (y - T) * M
    loss = 0.5 * (r * r).sum() / cnt
    dy = r / cnt
    dW3 = h2.T @ dy
    db3 = dy.sum(0)
    dh2 = dy @ W3.T
    dz2 = dh2 * (z2 > 0)
    dW2 = h1.T @ dz2
    db2 = dz2.sum(0)
    dh1 = dz2 @ W2.T
    dz1 = dh1 * (z1 > 0)
    dW1 = X.T @ dz1
    db1 = dz1.sum(0)
    return loss, [dW1, db1, dW2, db2, dW3, db3]


def adam_step(params, grads, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) in float64; returns new params, m, v."""
    out_p, out_m, out_v = [], [], []
    for p, g, mm, vv in zip(params, grads, m, v):
        mm = beta1 * mm + (1 - beta1) * g
        vv = beta2 * vv + (1 - beta2) * g * g
        mh = mm / (1 - beta1 ** t)
        vh = vv / (1 - beta2 ** t)
        out_p.append(p - lr * mh / (np.sqrt(vh) + eps))
        out_m.append(mm)
        out_v.append(vv)
    return out_p, out_m, out_v
