"""GPU parity at the benchmarked shapes, against the CPU oracle.

  production shape  base 2, 100 sequences with odd k < 2^32 (seed 2026, half
                    of them >= 2^31), n in [3*10^6, 10^7) pre-sieved to 2^20
                    (55 M terms), p in [2^50, 2^50 + 2^22) (about 121 k
                    primes): the reference split counter for counter
                    (primes, skipped, lists, order items, Montgomery products,
                    claims, duplicates, bitmap), the same with the reference's
                    signed-int k encoding (k >= 2^31 wraps, inputfile.cpp:366),
                    and the cost model's configuration sieved as four views in
                    one batch (bitmap, primes, claims)
  16 uneven jobs    bases {2,3,5,6,10}, 1-100 sequences, n-spans 10^4..10^7,
                    k < 2^31 (the descriptors of the benchmark's batch leg),
                    each with the cost model's 256-thread pick, all sixteen in
                    one batch over p in [2^48, 2^48 + 2^20): every job's
                    bitmap, prime count and claims equal the oracle's
"""
import os

import numpy as np
import pytest

import oracle as O
import rssieve_amd as R
from rssieve_amd import tuning
from util import cleared_terms, oracle_run

pytestmark = pytest.mark.gpu

NTHREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def production():
    import bench as B
    return B.make_workload()


def _oracle(base, ks, cs, nmin, nmax, bm, pmin, pmax, k31=False):
    return oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax, k31=k31, nthreads=NTHREADS, fac_cap=1 << 16)


def _check_factors(facs, ks, cs, base, k31=False):
    for f in facs:
        i = int(f["kidx"])
        k, c = ks[i], cs[i]
        if k31:  # the sequence the reference reports (cl_sieve.cpp:604-605)
            e = int(np.int32(np.uint32(k))) if c > 0 else int(np.int32(np.uint32((-k) & 0xFFFFFFFF)))
            k, c = abs(e), (1 if e > 0 else -1)
        assert R.verify_factor(int(f["p"]), k, int(f["n"]), c, base) in (1, -1)


@pytest.mark.parametrize("k31", [False, True], ids=["exact", "k31_compat"])
def test_production_shape_reference_split_matches_oracle(ctx, production, k31):
    ks, cs, nmin, nmax, bm, wps = production
    pmin, pmax = 2**50, 2**50 + 2**22
    print("production shape, reference split, k31_compat=%d" % k31, flush=True)
    prm = R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units)
    prm.k31_compat = 1 if k31 else 0
    prm.factor_cap = 1 << 16
    job = ctx.job(2, ks, cs, nmin, nmax, bm, wps, prm)
    job.sieve(pmin, pmax)
    g = job.counters().as_dict()
    gfac = job.drain(1 << 16)
    gbm = job.bitmap()
    job.close()
    o, ofac, obm = _oracle(2, ks, cs, nmin, nmax, bm, pmin, pmax, k31=k31)
    print("primes %d skipped %d lists %s claims %d dups %d products %d" % (
        o["primecount"], o["skipped"], [o["n_full"], o["n_even"], o["n_odd"]], o["claims"], o["dups"],
        o["mmul_work"]), flush=True)
    assert o["primecount"] > 100000
    assert g["seg_overflow"] == 0
    assert g["primecount"] == o["primecount"]
    assert g["skipped"] == o["skipped"]
    assert list(g["lists"]) == [o["n_full"], o["n_even"], o["n_odd"]]
    assert g["order_lists"] == o["order_lists"]
    assert g["work_mmul"] == o["mmul_work"]
    assert g["claims"] == o["claims"]
    assert g["dups"] == o["dups"]
    assert np.array_equal(gbm, obm)
    elim = cleared_terms(bm, gbm, wps, nmin)
    assert {(int(f["kidx"]), int(f["n"])) for f in gfac} == {(k, n) for (_, n, k) in ofac} == elim
    _check_factors(gfac, ks, cs, 2, k31)


def test_production_shape_cost_model_batched_matches_oracle(ctx, production):
    ks, cs, nmin, nmax, bm, wps = production
    pmin, pmax = 2**50, 2**50 + 2**22
    cands = ctx.tune_candidates(nmin, nmax, len(ks))
    idx = tuning.model_pick(ctx, cands, 2, len(ks), int(np.bitwise_count(bm).sum()), nmax - nmin + 1, 50.0)
    assert idx is not None, "the committed cost model does not match the candidate list"
    prm = cands[idx]
    prm.factor_cap = 1 << 16
    print("production shape, cost model pick %d: %s, four views in one batch" % (idx, prm.as_dict()), flush=True)
    job = ctx.job(2, ks, cs, nmin, nmax, bm, wps, prm)
    views = [job.view() for _ in range(3)]
    cut = [pmin + i * (pmax - pmin) // 4 for i in range(5)]
    R.sieve_batch([job] + views, cut[:4], cut[1:])
    g = job.counters().as_dict()
    gfac = job.drain(1 << 16)
    gbm = job.bitmap()
    for v in views:
        v.close()
    job.close()
    o, ofac, obm = _oracle(2, ks, cs, nmin, nmax, bm, pmin, pmax)
    assert g["seg_overflow"] == 0
    assert g["primecount"] == o["primecount"]
    assert g["skipped"] == o["skipped"]
    assert g["claims"] == o["claims"]
    assert np.array_equal(gbm, obm)
    assert {(int(f["kidx"]), int(f["n"])) for f in gfac} == {(k, n) for (_, n, k) in ofac}
    _check_factors(gfac, ks, cs, 2)


def test_sixteen_uneven_jobs_in_one_batch_match_oracle(ctx):
    descs = tuning.draw_descriptors(16, 16, log2p_range=(48.0, 48.0))
    model = tuning.load_model(ctx)
    p0, p1 = 2**48, 2**48 + 2**20
    cache, jobs, inputs = {}, [], []
    for d in descs:
        nmax_d = d["nmin"] + d["span"] - 1
        prm = R.reference_params(d["nmin"], nmax_d, d["K"], 65536, ctx.compute_units)
        job, nterms, nmin, nmax = tuning.synth_job(ctx, d, prm, cache)
        cands = ctx.tune_candidates(nmin, nmax, d["K"])
        assert len(cands) == model.n_out
        x = tuning.features(d["base"], d["K"], nterms, d["span"], 48.0)[None, :]
        idx = int(model.choose(x, np.array([[c.threads == 256 for c in cands]]))[0])
        job.reconfigure(cands[idx])
        jobs.append(job)
        bits, _ = cache[(d["seed"], d["K"], nmin, nmax)]
        inputs.append((d, nmin, nmax, bits.copy(), idx))
    print("16 jobs (base, K, span, candidate):", [(d["base"], d["K"], d["span"], idx)
                                                  for (d, _, _, _, idx) in inputs], flush=True)
    R.sieve_batch(jobs, [p0] * len(jobs), [p1] * len(jobs))
    for job, (d, nmin, nmax, bits, idx) in zip(jobs, inputs):
        g = job.counters().as_dict()
        gfac = job.drain()
        gbm = job.bitmap()
        job.close()
        ks, cs = tuning.synth_sequences(d)
        o, ofac, obm = _oracle(d["base"], ks, cs, nmin, nmax, bits, p0, p1)
        print("job base %d K %d span %d: primes %d claims %d" % (d["base"], d["K"], d["span"], o["primecount"],
                                                                 o["claims"]), flush=True)
        assert g["primecount"] == o["primecount"]
        assert g["skipped"] == o["skipped"]
        assert g["claims"] == o["claims"]
        assert np.array_equal(gbm, obm)
        assert {(int(f["kidx"]), int(f["n"])) for f in gfac} == {(k, n) for (_, n, k) in ofac}
        _check_factors(gfac, ks, cs, d["base"])
