"""GPU tests of the configuration search: the MLP's HIP forward/backward/Adam
kernels against the CPU reference (oracle/mlp_ref.py, pinned by finite
differences in test_mlp_ref.py), and the autotuner's contract (every
candidate timed on the same sample, the job's bitmap and counters restored,
the chosen configuration sieving exactly like the oracle).

Tolerance (fp32 kernels vs float64 reference): |a - b| <= 1e-4 * max(1, |b|)
elementwise for outputs, loss and gradients."""
import numpy as np
import pytest

import mlp_ref as MR
import rssieve_amd as R
from util import full_bitmap, oracle_run

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _close(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= TOL * np.maximum(1.0, np.abs(b)))


@pytest.mark.parametrize("shape", [(5, 64, 25), (5, 16, 3), (7, 33, 70)])
def test_mlp_forward_and_gradients_match_reference(ctx, shape):
    n_in, H, C = shape
    rng = np.random.default_rng(sum(shape))
    m = ctx.mlp(n_in, H, C, seed=3)
    params = [p + 0.05 * rng.standard_normal(p.shape).astype(np.float32) for p in MR.init(n_in, H, C, seed=4)]
    m.set_params(params)
    got = m.get_params()
    assert all(np.array_equal(a, b) for a, b in zip(got, params))
    for n in (1, 37, 300):
        X = rng.standard_normal((n, n_in)).astype(np.float32)
        T = rng.standard_normal((n, C)).astype(np.float32)
        M = (rng.random((n, C)) < 0.7).astype(np.float32)
        y_ref, _ = MR.forward(params, X)
        assert _close(m.forward(X), y_ref)
        loss, grads = m.loss_grad(X, T, M)
        loss_ref, grads_ref = MR.loss_and_grads(params, X, T, M)
        assert abs(loss - loss_ref) <= TOL * max(1.0, abs(loss_ref))
        for g, gr in zip(grads, grads_ref):
            assert _close(g, gr)
    m.close()


def test_mlp_adam_step_matches_reference(ctx):
    rng = np.random.default_rng(9)
    m = ctx.mlp(5, 64, 8, seed=1)
    params = MR.init(5, 64, 8, seed=2)
    m.set_params(params)
    X = rng.standard_normal((64, 5)).astype(np.float32)
    T = rng.standard_normal((64, 8)).astype(np.float32)
    p = [x.astype(np.float64) for x in params]
    mm = [np.zeros_like(x) for x in p]
    vv = [np.zeros_like(x) for x in p]
    for t in range(1, 4):
        _, g = m.loss_grad(X, T)
        m.adam_step(1e-2)
        _, gr = MR.loss_and_grads(p, X, T)
        p, mm, vv = MR.adam_step(p, gr, mm, vv, t, lr=1e-2)
    for a, b in zip(m.get_params(), p):
        assert np.max(np.abs(a - b)) < 2e-4
    m.close()


def test_mlp_training_fits_a_smooth_target(ctx):
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, size=(2048, 5)).astype(np.float32)
    W = rng.standard_normal((5, 6))
    T = (np.sin(X @ W) + 0.5 * (X[:, :1] ** 2)).astype(np.float32)
    m = ctx.mlp(5, 64, 6, seed=7)
    l0, _ = m.loss_grad(X, T)
    l1 = m.train(X, T, epochs=60, batch=128, lr=3e-3, seed=11)
    assert l1 < 0.1 * l0, (l0, l1)
    m.close()


def test_autotune_restores_state_and_best_config_is_exact(ctx):
    base, nmin, nmax = 2, 200000, 260000
    ks = [3, 5, 7, 11, 13, 15, 17, 19, 21, 23, 25, 27]
    cs = [1, -1] * 6
    bm, wps = full_bitmap(len(ks), nmin, nmax, density=0.5, seed=3)
    prm = R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units)
    job = ctx.job(base, ks, cs, nmin, nmax, bm, wps, prm)
    cands = ctx.tune_candidates(nmin, nmax, len(ks))
    assert cands[0].algo == R.ALGO_THREE_LISTS and cands[0].Q == prm.Q
    assert len(cands) >= 20
    pmin, width = 10**9, 3 * 10**6
    print("autotune over", len(cands), "candidates", flush=True)
    best, secs, primes = job.autotune(pmin, width, cands)
    assert 0 <= best < len(cands)
    ran = [c.threads != 0 for c in cands]  # slots this n-range cannot use are marked threads = 0
    assert all((s > 0) == r for s, r in zip(secs, ran))
    assert len({pr for pr, r in zip(primes, ran) if r}) == 1 and primes[0] > 0
    assert secs[best] == min(s for s, r in zip(secs, ran) if r)
    # the tuning left no trace
    assert np.array_equal(job.bitmap(), bm)
    c = job.counters()
    assert c.primecount == 0 and c.claims == 0 and c.dups == 0
    # the job now runs the best configuration, with the oracle's eliminations
    job.sieve(pmin, pmin + width)
    gbm = job.bitmap()
    facs = job.drain()
    ocnt, ofac, obm = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmin + width)
    assert np.array_equal(gbm, obm)
    assert len(facs) == ocnt["claims"]
    job.close()
