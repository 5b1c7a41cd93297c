"""Pins the CPU reference of the cost-model MLP (oracle/mlp_ref.py) by
central finite differences and by the closed form of one Adam step."""
import numpy as np

import mlp_ref as M


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    params = [p.astype(np.float64) for p in M.init(5, 16, 7, seed=1)]
    params = [p + 0.01 * rng.standard_normal(p.shape) for p in params]  # nonzero biases too
    X = rng.standard_normal((13, 5))
    T = rng.standard_normal((13, 7))
    Msk = (rng.random((13, 7)) < 0.8).astype(np.float64)
    loss, grads = M.loss_and_grads(params, X, T, Msk)
    eps = 1e-6
    for pi, p in enumerate(params):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, size=min(flat.size, 25), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = M.loss_and_grads(params, X, T, Msk)
            flat[idx] = old - eps
            lm, _ = M.loss_and_grads(params, X, T, Msk)
            flat[idx] = old
            num = (lp - lm) / (2 * eps)
            assert abs(num - grads[pi].reshape(-1)[idx]) <= 1e-6 + 1e-5 * abs(num), (pi, idx)


def test_masked_entries_do_not_contribute():
    rng = np.random.default_rng(1)
    params = M.init(5, 8, 4, seed=2)
    X = rng.standard_normal((6, 5))
    T = rng.standard_normal((6, 4))
    Msk = np.ones((6, 4))
    Msk[:, 2] = 0
    T2 = T.copy()
    T2[:, 2] += 100.0
    l1, g1 = M.loss_and_grads(params, X, T, Msk)
    l2, g2 = M.loss_and_grads(params, X, T2, Msk)
    assert l1 == l2 and all(np.array_equal(a, b) for a, b in zip(g1, g2))


def test_adam_step_closed_form():
    p = [np.array([1.0, -2.0])]
    g = [np.array([0.5, -0.25])]
    m = [np.zeros(2)]
    v = [np.zeros(2)]
    p1, m1, v1 = M.adam_step(p, g, m, v, t=1, lr=0.1)
    # first step: m_hat = g, v_hat = g^2, update = lr * sign(g) (up to eps)
    assert np.allclose(p1[0], p[0] - 0.1 * np.sign(g[0]), atol=1e-7)
