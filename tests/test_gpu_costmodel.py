"""What the learned cost model gains over the configuration rule without it.

Data: tests/golden/tune_records.npz, autotuner timings of synthetic jobs
(written by tests/golden/make_tune_records.py).  The MLP (5 -> 64 -> 64 ->
#candidates, fp32, the library's HIP kernels) is trained on the first rows and
judged on the last 256: its regret (time of its pick over the measured best)
against the reference split's (build_giant_kernels, candidate 0), the rule a
job gets without the model, and against the best single candidate of the
training rows (a non-learned rule fitted to the same data)."""
import os

import numpy as np
import pytest

import rssieve_amd as R
from rssieve_amd import tuning

pytestmark = pytest.mark.gpu

RECORDS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tune_records.npz")


@pytest.mark.skipif(not os.path.isfile(RECORDS), reason="no recorded autotuner timings")
def test_cost_model_beats_reference_rule(ctx):
    z = np.load(RECORDS, allow_pickle=False)
    X, Y, V = z["X"], z["Y"], z["V"]
    n_test = 256
    ntr = len(X) - n_test
    assert ntr >= 256
    m = tuning.CostModel(ctx, Y.shape[1], seed=5)
    m.fit(X[:ntr], Y[:ntr], V[:ntr])
    choice = m.choose(X[ntr:], V[ntr:])
    r_model = tuning.regret(Y[ntr:], V[ntr:], choice)
    r_ref = tuning.regret(Y[ntr:], V[ntr:], np.zeros(n_test, np.int64))
    fixed = tuning.best_fixed(Y[:ntr], V[:ntr])
    r_fix = tuning.regret(Y[ntr:], V[ntr:], np.full(n_test, fixed, np.int64))
    gain = float(np.exp(np.mean(np.log((1 + r_ref) / (1 + r_model)))))
    print("held-out regret: model mean %.3f median %.3f | reference rule mean %.3f median %.3f | best fixed (%d) "
          "mean %.3f | geomean gain over reference %.2fx" % (r_model.mean(), np.median(r_model), r_ref.mean(),
                                                              np.median(r_ref), fixed, r_fix.mean(), gain), flush=True)
    assert r_model.mean() < r_ref.mean()
    assert r_model.mean() < r_fix.mean()
    assert gain > 1.1
    # the untrained model falls back to the reference rule
    fresh = tuning.CostModel(ctx, Y.shape[1])
    assert np.all(fresh.choose(X[ntr:]) == 0)


SPLIT_RECORDS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "split_records.npz")


@pytest.mark.skipif(not os.path.isfile(SPLIT_RECORDS), reason="no recorded batch-split timings")
def test_split_head_regret(ctx):
    """The batch-split head (views per batched round and primes per segment,
    rs_autotune_split timings of 384 synthetic jobs, written by
    tests/golden/make_split_records.py): trained on the first 256 jobs, its
    pick on the last 128 against the measured best.  It must beat one view per
    round (no batch split) by far and stay within 1 % of the best on average;
    the fixed splits are printed beside it (larger launches win for nearly
    every job, so the largest option is close to the best as well)."""
    import json
    z = np.load(SPLIT_RECORDS, allow_pickle=False)
    X, Y, V = z["X"], z["Y"], z["V"]
    options = [tuple(o) for o in json.loads(str(z["options"]))]
    ntr = 256
    m = tuning.SplitModel(ctx, options, seed=3)
    m.fit(X[:ntr], Y[:ntr], V[:ntr])
    choice = m.choose(X[ntr:], V[ntr:])
    r_model = tuning.regret(Y[ntr:], V[ntr:], choice)
    fixed = {o: tuning.regret(Y[ntr:], V[ntr:], np.full(len(choice), i, np.int64)).mean()
             for i, o in enumerate(options)}
    print("held-out split regret: model mean %.4f | fixed %s" % (
        r_model.mean(), ", ".join("%dx%d %.4f" % (v, sp, r) for (v, sp), r in fixed.items())), flush=True)
    assert r_model.mean() < fixed[(1, 65536)] / 5
    assert r_model.mean() < 0.01


def test_model_rank_through_the_c_abi(ctx):
    """rs_model_rank (the committed model through the library): the eligible
    candidates in order of prediction, the first being rs_model_choose's pick;
    a thread filter leaves only candidates of that workgroup size, and a
    candidate list of the wrong length is refused."""
    import bench
    cm = tuning.CModel(ctx)
    try:
        for nmin, nmax, K, log2p in ((3000000, 9999999, 100, 50.0), (10**6, 2**24 - 1, 100, 63.0),
                                     (100000, 199999, 10, 30.0)):
            cands = ctx.tune_candidates(nmin, nmax, K)
            for threads_only in (0, 256):
                first = cm.choose(2, K, 10**7, nmax - nmin + 1, log2p, cands, threads_only)
                top = cm.rank(2, K, 10**7, nmax - nmin + 1, log2p, cands, 3, threads_only)
                assert top[0] == first and len(set(top)) == len(top) == 3
                assert all(cands[i].threads and (not threads_only or cands[i].threads == threads_only) for i in top)
                assert cm.rank(2, K, 10**7, nmax - nmin + 1, log2p, cands, 1, threads_only) == [first]
        with pytest.raises(R.RSError):
            cm.rank(2, 100, 10**7, 7 * 10**6, 50.0, cands[:-1], 3)
        want = cm.rank(2, 100, 10**7, 7 * 10**6, 50.0, cands, 3)
    finally:
        cm.close()
    assert bench.model_choice(ctx, 2, 100, 10**7, 7 * 10**6, 50.0, cands)[0] == want
