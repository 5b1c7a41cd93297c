"""rs_sieve_noalloc (the allocation-free sieve entry point) against rs_sieve,
through the C ABI, on two calls in a row, the second with a smaller job.

Both jobs run with a one-entry positive queue, so nearly every BSGS round
overflows and is redone from the overflow-recovery tables: the tables that
rs_sieve_noalloc keeps in its static device scratch are used on every item.
The bar is bit-exact: every counter (primes, claims, duplicates, lists,
order items, products, probes), the bitmap and the eliminated (k, n) set.
A shape whose recovery tables do not fit the static scratch is refused.
"""
import numpy as np
import pytest

import rssieve_amd as R
from util import full_bitmap

pytestmark = pytest.mark.gpu


def _draw(base, K, seed):
    rng = np.random.default_rng(seed)
    ks, cs = [], []
    while len(ks) < K:
        k = int(rng.integers(3, 10**6))
        if (k % 2 == 0) == (base % 2 == 0) or np.gcd(k, base) != 1 or k in ks:
            continue
        ks.append(k)
        cs.append(1 if len(cs) % 2 else -1)
    return ks, cs


def _sieve(ctx, noalloc, base, ks, cs, nmin, nmax, bm, wps, prm, pmin, pmax):
    job = ctx.job(base, ks, cs, nmin, nmax, bm, wps, prm)
    if noalloc:
        job.sieve_noalloc(pmin, pmax)
    else:
        job.sieve(pmin, pmax)
    cnt = job.counters().as_dict()
    facs = job.drain()
    got = job.bitmap()
    job.close()
    return cnt, {(int(f["kidx"]), int(f["n"])) for f in facs}, got


# base, K, nmin, nmax, pmin, pmax, seed, algo, Q (entries), hsize, threads
FIRST = (2, 10, 100000, 199999, 10**9, 10**9 + 3 * 10**6, 41, R.ALGO_THREE_LISTS, 0, 0, 256)
SECOND = (3, 4, 20000, 50000, 10**7, 10**7 + 5 * 10**5, 42, R.ALGO_CLASSES, 2048, 4096, 256)


@pytest.mark.parametrize("qcap", [1, 0], ids=["queue1", "queue_max"])
def test_noalloc_matches_rs_sieve_on_two_calls(ctx, qcap):
    for case in (FIRST, SECOND):
        base, K, nmin, nmax, pmin, pmax, seed, algo, Q, hsize, threads = case
        print("no-alloc case", case, "queue_cap", qcap, flush=True)
        ks, cs = _draw(base, K, seed)
        bm, wps = full_bitmap(K, nmin, nmax, density=0.6, seed=seed)
        prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
        if algo != R.ALGO_THREE_LISTS:
            prm.algo, prm.Q, prm.hsize, prm.threads = algo, Q, hsize, threads
        prm.queue_cap = qcap
        want = _sieve(ctx, False, base, ks, cs, nmin, nmax, bm, wps, prm, pmin, pmax)
        got = _sieve(ctx, True, base, ks, cs, nmin, nmax, bm, wps, prm, pmin, pmax)
        assert got[0] == want[0]
        assert got[1] == want[1]
        assert np.array_equal(got[2], want[2])
        assert want[0]["claims"] > 0 and want[0]["seg_overflow"] == 0


def test_noalloc_refuses_tables_beyond_the_static_scratch(ctx):
    # residue-class tables of 16384 entries at 128 threads: two workgroups per
    # CU (the filter takes 64 KiB of LDS) with 128 KiB recovery tables each
    base, K, nmin, nmax = 2, 2, 1000, 400000
    ks, cs = _draw(base, K, 43)
    bm, wps = full_bitmap(K, nmin, nmax)
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    prm.algo, prm.Q, prm.hsize, prm.threads = R.ALGO_CLASSES, 16384, 16384, 128
    job = ctx.job(base, ks, cs, nmin, nmax, bm, wps, prm)
    with pytest.raises(R.RSError) as e:
        job.sieve_noalloc(10**9, 10**9 + 10**5)
    job.close()
    assert e.value.code == R.RS_E_ARG and "static scratch" in str(e.value)
