"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs, at sizes the oracle finishes in seconds.

Bar: bit-exact.  The cleared bitmap, the prime count, the number of
successful claims and of duplicate claims must be identical; the set of
eliminated (k, n) terms must be identical; every factor the GPU reports must
divide its term exactly.  Which prime is credited for a term with two prime
factors in range is decided by an atomic race in the reference and here
(giant.cl:36), so factor lines are compared on the terms whose credited prime
agrees and checked exactly on the rest.
"""
import numpy as np
import pytest

import oracle as O
import rssieve_amd as R
from util import cleared_terms, full_bitmap, oracle_run

pytestmark = pytest.mark.gpu


def _primes_oracle(lo, hi):
    return np.sort(O.primes_in_range(lo, hi, seg=1 << 22))


@pytest.mark.parametrize("lo,hi", [
    (3, 20000),
    (100, 114),
    (113, 1000),
    (10**9, 10**9 + 3 * 10**6),
    (2**50, 2**50 + 2 * 10**6),
    (2**63 - 10**6, 2**63 + 10**6),
    (2**64 - 3 * 10**6, 2**64 - 1),
])
def test_primegen_matches_oracle(ctx, lo, hi):
    print("primegen case", lo, hi, flush=True)
    got = np.sort(ctx.primegen(lo, hi))
    want = _primes_oracle(lo, hi)
    assert got.size == want.size
    assert np.array_equal(got, want)


def _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, params):
    job = ctx.job(base, ks, cs, nmin, nmax, bm, wps, params)
    job.sieve(pmin, pmax)
    cnt = job.counters().as_dict()
    facs = job.drain()
    bm_out = job.bitmap()
    job.close()
    return cnt, facs, bm_out


def _compare(base, ks, cs, nmin, nmax, wps, bm_in, gpu, orc, k31=False):
    gcnt, gfac, gbm = gpu
    ocnt, ofac, obm = orc
    assert gcnt["seg_overflow"] == 0
    assert gcnt["primecount"] == ocnt["primecount"]
    assert gcnt["skipped"] == ocnt["skipped"]
    assert list(gcnt["lists"]) == [ocnt["n_full"], ocnt["n_even"], ocnt["n_odd"]]
    assert gcnt["order_lists"] == ocnt["order_lists"]
    # Montgomery products of the reference's three tables: baby steps up to
    # the table size (or the repeat of the order shortcut) plus M * (targets
    # + 1) giant-step products per list
    assert gcnt["work_mmul"] == ocnt["mmul_work"]
    assert gcnt["claims"] == ocnt["claims"]
    assert gcnt["dups"] == ocnt["dups"]
    assert np.array_equal(gbm, obm)
    elim = cleared_terms(bm_in, gbm, wps, nmin)
    gterms = {(int(f["kidx"]), int(f["n"])) for f in gfac}
    oterms = {(k, n) for (_, n, k) in ofac}
    assert gterms == oterms == elim
    assert len(gfac) == len(ofac) == gcnt["claims"]
    opmap = {(k, n): p for (p, n, k) in ofac}
    for f in gfac:
        k, n, p = int(f["kidx"]), int(f["n"]), int(f["p"])
        if opmap[(k, n)] != p:
            kk, cc = (ks[k], cs[k])
            if k31:
                e = int(np.int32(np.uint32(kk))) if cc > 0 else int(np.int32(np.uint32((-int(kk)) & 0xFFFFFFFF)))
                kk, cc = abs(e), (1 if e > 0 else -1)
            assert pow(base, n, p) * kk % p == (p - cc) % p or R.verify_factor(p, kk, n, cc, base) == -1


CASES = [
    # name, base, K, nmin, nmax, pmin, pmax, seed, density
    ("b2_k10_1e9", 2, 10, 100000, 110000, 10**9, 10**9 + 2 * 10**6, 1, 1.0),
    ("b2_k6_lowp", 2, 6, 2000, 6000, 200, 60000, 2, 1.0),
    ("b3_k4", 3, 4, 1000, 9000, 10**6, 10**6 + 2 * 10**5, 3, 1.0),
    ("b5_k5", 5, 5, 5000, 20000, 10**7, 10**7 + 2 * 10**5, 4, 1.0),
    ("b6_k3", 6, 3, 3000, 30000, 10**7, 10**7 + 10**5, 5, 1.0),
    ("b10_k7", 10, 7, 800, 12000, 10**5, 10**5 + 4 * 10**5, 6, 1.0),
    ("b2_k20_2e40", 2, 20, 300000, 420000, 2**40, 2**40 + 2 * 10**5, 7, 0.5),
    ("b2_k100_2e32", 2, 100, 50000, 60000, 2**32 - 10**5, 2**32 + 10**5, 8, 0.3),
    ("b7_k2_2e63", 7, 2, 10, 70000, 2**63 - 10**4, 2**63 + 10**4, 9, 1.0),
    ("b2_k5_top", 2, 5, 1000, 40000, 2**64 - 2 * 10**4, 2**64 - 1, 10, 1.0),
    # p below / above 2^51: binary64 steps, then integer Montgomery steps
    ("b2_k12_2e51", 2, 12, 60000, 90000, 2**51 - 10**5, 2**51 + 10**5, 11, 0.7),
    ("b3_k8_2e50", 3, 8, 20000, 50000, 2**50 - 10**5, 2**50 + 10**5, 12, 1.0),
]


def _draw(base, K, seed, k_hi=10**6):
    """K seeded k coprime to base, odd for even bases and even for odd bases
    (so the terms are odd), with c alternating -1, +1."""
    rng = np.random.default_rng(seed)
    ks, cs = [], []
    while len(ks) < K:
        k = int(rng.integers(3, k_hi))
        if (k % 2 == 0) == (base % 2 == 0) or np.gcd(k, base) != 1 or k in ks:
            continue
        ks.append(k)
        cs.append(1 if len(cs) % 2 else -1)
    return ks, cs


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_sieve_matches_oracle(ctx, case):
    name, base, K, nmin, nmax, pmin, pmax, seed, dens = case
    print("sieve case", case, flush=True)
    ks, cs = _draw(base, K, seed)
    bm, wps = full_bitmap(K, nmin, nmax, density=dens, seed=seed)
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
    _compare(base, ks, cs, nmin, nmax, wps, bm, gpu, orc)


@pytest.mark.parametrize("k31", [False, True])
def test_sieve_k_above_2_31(ctx, k31):
    """k >= 2^31: exact by default, the reference's signed wrap with k31_compat."""
    base, nmin, nmax = 2, 20000, 60000
    rng = np.random.default_rng(11)
    ks = [int(x) | 1 for x in rng.integers(2**31, 2**32 - 1, size=8)]
    cs = [1, -1] * 4
    bm, wps = full_bitmap(len(ks), nmin, nmax)
    prm = R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units)
    prm.k31_compat = 1 if k31 else 0
    pmin, pmax = 10**8, 10**8 + 10**6
    print("k31 case", k31, flush=True)
    gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax, k31=k31)
    _compare(base, ks, cs, nmin, nmax, wps, bm, gpu, orc, k31=k31)


def test_sieve_k_2_31_compat_wrap(ctx):
    """k = 2^31 under the reference's signed-int encoding: the entry is -2^31
    whatever c is, and its magnitude reaches the setup sign-extended to 64
    bits (setup.cl:143).  Base 3 (even k), p below and above 2^51."""
    base, nmin, nmax = 3, 20000, 60000
    ks = [2**31, 2**31, 2**31 + 2, 2**32 - 2]
    cs = [1, -1, 1, -1]
    bm, wps = full_bitmap(len(ks), nmin, nmax)
    prm = R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units)
    prm.k31_compat = 1
    for pmin, pmax in ((10**8, 10**8 + 10**6), (2**51 - 10**5, 2**51 + 10**5)):
        print("k = 2^31 compat case", pmin, pmax, flush=True)
        gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
        orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax, k31=True)
        _compare(base, ks, cs, nmin, nmax, wps, bm, gpu, orc, k31=True)


OVERFLOW_CASES = [c for c in CASES if c[0] in ("b2_k10_1e9", "b2_k20_2e40", "b7_k2_2e63", "b2_k12_2e51")]


@pytest.mark.parametrize("case", OVERFLOW_CASES, ids=[c[0] for c in OVERFLOW_CASES])
@pytest.mark.parametrize("qcap", [1, 5])
def test_positive_queue_overflow_is_recovered(ctx, case, qcap):
    """A BSGS workgroup queues filter positives and resolves them at round
    boundaries; with a queue of 1 or 5 entries nearly every round overflows
    and is redone by direct lookups.  Results, duplicate count included, stay
    the oracle's (binary64 and integer paths, both table shapes)."""
    name, base, K, nmin, nmax, pmin, pmax, seed, dens = case
    print("overflow case", case, qcap, flush=True)
    ks, cs = _draw(base, K, seed)
    bm, wps = full_bitmap(K, nmin, nmax, density=dens, seed=seed)
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    prm.queue_cap = qcap
    gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
    _compare(base, ks, cs, nmin, nmax, wps, bm, gpu, orc)
    for algo in (R.ALGO_MERGED, R.ALGO_CLASSES):
        mprm = merged_params(ctx, nmin, nmax, K, 2048, 4096, 512, algo=algo)
        mprm.queue_cap = qcap
        gcnt, gfac, gbm = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, mprm)
        assert gcnt["claims"] == orc[0]["claims"]
        assert np.array_equal(gbm, orc[2])


def merged_params(ctx, nmin, nmax, K, entries=4096, hsize=8192, threads=256, algo=R.ALGO_MERGED):
    """Parameters of the one-list paths: the merged table of 'entries' even
    exponent offsets (span 2*entries), or residue-class tables of at most
    'entries' entries."""
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    prm.algo = algo
    prm.Q = 2 * entries if algo == R.ALGO_MERGED else entries
    prm.hsize = hsize
    prm.threads = threads
    return prm


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("shape", [(1024, 2048, 128), (4096, 8192, 256), (8192, 16384, 512)],
                         ids=["e1024", "e4096", "e8192"])
@pytest.mark.parametrize("algo", [R.ALGO_MERGED, R.ALGO_CLASSES], ids=["merged", "classes"])
def test_merged_table_same_eliminations(ctx, case, shape, algo):
    """The one-list BSGS shapes eliminate exactly the terms the
    reference-shaped sieve (and the oracle) eliminates: the merged table, and
    the residue-class tables, whose stride and length change from prime to
    prime (p below 1024 keep the parity split).  Their duplicate counts
    follow their own search windows, so they are not compared."""
    name, base, K, nmin, nmax, pmin, pmax, seed, dens = case
    print("one-list case", case, shape, algo, flush=True)
    ks, cs = _draw(base, K, seed)
    bm, wps = full_bitmap(K, nmin, nmax, density=dens, seed=seed)
    prm = merged_params(ctx, nmin, nmax, K, *shape, algo=algo)
    gcnt, gfac, gbm = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    ocnt, ofac, obm = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
    assert gcnt["primecount"] == ocnt["primecount"]
    assert gcnt["skipped"] == ocnt["skipped"]
    assert gcnt["claims"] == ocnt["claims"]
    assert np.array_equal(gbm, obm)
    assert {(int(f["kidx"]), int(f["n"])) for f in gfac} == {(k, n) for (_, n, k) in ofac}
    for f in gfac:
        i = int(f["kidx"])
        assert R.verify_factor(int(f["p"]), ks[i], int(f["n"]), cs[i], base) in (1, -1)


def test_batch_of_uneven_jobs_matches_oracle(ctx):
    """Several independent jobs (bases, sequence counts, n-spans, p-ranges and
    both BSGS shapes) in one batch: one persistent BSGS launch per round,
    replayed from a HIP graph.  Each job's result equals the oracle's."""
    specs = [
        # base, K, nmin, nmax, pmin, pmax, seed, algo, entries, hsize
        (2, 10, 100000, 110000, 10**9, 10**9 + 10**6, 21, R.ALGO_THREE_LISTS, 0, 0),
        (3, 1, 2000, 40000, 10**7, 10**7 + 3 * 10**5, 22, R.ALGO_MERGED, 2048, 8192),
        (5, 7, 10000, 30000, 2**33, 2**33 + 2 * 10**5, 23, R.ALGO_THREE_LISTS, 0, 0),
        (6, 3, 500, 2500, 300, 70000, 24, R.ALGO_MERGED, 1024, 4096),
        (10, 25, 40000, 60000, 10**6, 10**6 + 4 * 10**5, 25, R.ALGO_MERGED, 4096, 16384),
        (2, 100, 3000000, 3010000, 2**48, 2**48 + 2 * 10**5, 26, R.ALGO_THREE_LISTS, 0, 0),
        (6, 40, 20000, 90000, 2**40, 2**40 + 2 * 10**5, 27, R.ALGO_CLASSES, 4096, 8192),
        (7, 9, 100, 60000, 2**62, 2**62 + 10**5, 28, R.ALGO_CLASSES, 2048, 2048),
    ]
    jobs, inputs = [], []
    for (base, K, nmin, nmax, pmin, pmax, seed, algo, ent, hs) in specs:
        ks, cs = _draw(base, K, seed)
        bm, wps = full_bitmap(K, nmin, nmax, density=0.5, seed=seed)
        prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
        if algo == R.ALGO_MERGED:
            prm.algo, prm.Q, prm.hsize = algo, 2 * ent, hs
        elif algo == R.ALGO_CLASSES:
            prm.algo, prm.Q, prm.hsize = algo, ent, hs
        jobs.append(ctx.job(base, ks, cs, nmin, nmax, bm, wps, prm))
        inputs.append((base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, algo))
    print("batch of", len(jobs), flush=True)
    R.sieve_batch(jobs, [x[7] for x in inputs], [x[8] for x in inputs])
    for job, (base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, algo) in zip(jobs, inputs):
        cnt = job.counters().as_dict()
        facs = job.drain()
        gbm = job.bitmap()
        job.close()
        ocnt, ofac, obm = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
        assert cnt["primecount"] == ocnt["primecount"]
        assert cnt["claims"] == ocnt["claims"]
        if algo == R.ALGO_THREE_LISTS:
            assert cnt["dups"] == ocnt["dups"]
        assert np.array_equal(gbm, obm)
        assert {(int(f["kidx"]), int(f["n"])) for f in facs} == {(k, n) for (_, n, k) in ofac}


def _high_p_data(seed=63, K=100, nmin=10**6, nmax=2**24 - 1, and_words=5):
    """The high-p production shape: base 5, K sequences with k in [2, 2^32)
    coprime to 5 (even and odd k), c alternating -1, +1, n in [10^6, 2^24),
    and a bitmap keeping about 1/32 of the terms (the AND of five random
    words per word), the share a pre-sieve to 2^20 leaves."""
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < K:
        k = int(rng.integers(2, 2**32))
        if k % 5 == 0 or k in ks:
            continue
        ks.append(k)
    cs = [1 if i % 2 else -1 for i in range(K)]
    wps = (nmax - nmin + 1 + 31) // 32
    bm = rng.integers(0, 2**32, size=K * wps, dtype=np.uint32)
    for _ in range(and_words - 1):
        bm &= rng.integers(0, 2**32, size=K * wps, dtype=np.uint32)
    tail = (nmax - nmin + 1) % 32
    if tail:  # bits past nmax stay clear
        bm.reshape(K, wps)[:, -1] &= np.uint32((1 << tail) - 1)
    return ks, cs, nmin, nmax, bm, wps


def test_high_p_production_shape_batched_and_tuned(ctx):
    """b = 5, 100 sequences, n in [10^6, 2^24), p from 2^63 (integer
    Montgomery path near the top of the 64-bit range): the reference-shaped
    sieve equals the oracle counter for counter, and the batched path (two
    views of the job in one batch) with the cost model's configuration (the
    largest residue-class table without a model for this candidate list)
    leaves the same bitmap."""
    base = 5
    ks, cs, nmin, nmax, bm, wps = _high_p_data()
    K = len(ks)
    pmin, pmax = 2**63, 2**63 + 2 * 10**4
    print("high-p case", base, K, nmin, nmax, pmin, pmax, flush=True)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax, nthreads=8)
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    gcnt, gfac, gbm = gpu
    ocnt, ofac, obm = orc
    assert gcnt["primecount"] == ocnt["primecount"] > 400
    assert gcnt["skipped"] == ocnt["skipped"]
    assert list(gcnt["lists"]) == [ocnt["n_full"], ocnt["n_even"], ocnt["n_odd"]]
    assert gcnt["order_lists"] == ocnt["order_lists"]
    assert gcnt["claims"] == ocnt["claims"] and gcnt["dups"] == ocnt["dups"]
    assert np.array_equal(gbm, obm)

    from rssieve_amd import tuning
    cands = ctx.tune_candidates(nmin, nmax, K)
    idx = tuning.model_pick(ctx, cands, base, K, int(np.bitwise_count(bm).sum()), nmax - nmin + 1, 63.0)
    if idx is None:
        idx = max((i for i, c in enumerate(cands) if c.algo == R.ALGO_CLASSES and c.threads), key=lambda i: cands[i].Q)
    print("tuned candidate", idx, cands[idx].as_dict(), flush=True)
    job = ctx.job(base, ks, cs, nmin, nmax, bm, wps, cands[idx])
    view = job.view()
    mid = pmin + (pmax - pmin) // 3
    R.sieve_batch([job, view], [pmin, mid], [mid, pmax])
    cnt = job.counters().as_dict()
    facs = job.drain()
    tbm = job.bitmap()
    view.close()
    job.close()
    assert cnt["primecount"] == ocnt["primecount"]
    assert cnt["skipped"] == ocnt["skipped"]
    assert cnt["claims"] == ocnt["claims"]
    assert np.array_equal(tbm, obm)
    for f in facs:
        i = int(f["kidx"])
        assert R.verify_factor(int(f["p"]), ks[i], int(f["n"]), cs[i], base) == 1


@pytest.mark.gpu
def test_small_case_whole_range_matches_oracle(ctx):
    """The reference's CPU-sized case (b=2, ten odd k below 10^6, n in [10^5, 2*10^5)
    pre-sieved to 2^16) over its whole p-range [10^9, 1.1*10^9), 4.81 M primes: the reference
    split and the cost model's pick both leave the oracle's bitmap and count its primes; the
    reference split also makes its claims."""
    import os
    import bench as B
    from rssieve_amd import tuning
    ks, cs, nmin, nmax, bm, wps = B._small_case()
    p0, p1 = 10**9, 11 * 10**8
    ob = bm.copy()
    cnt, _, _ = O.sieve(O.make_job(2, ks, cs, nmin, nmax), p0, p1, ob, fac_cap=1 << 12,
                        nthreads=max(1, min(16, os.cpu_count() or 1)))
    assert cnt.primecount == 4814969 and cnt.claims > 0
    model = tuning.load_model(ctx)
    cands = ctx.tune_candidates(nmin, nmax, len(ks))
    x = tuning.features(2, len(ks), int(np.bitwise_count(bm).sum()), nmax - nmin + 1, np.log2(float(p0)))
    idx = int(model.choose(x[None, :], np.array([[c.threads != 0 for c in cands]]))[0])
    for prm in (R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units), cands[idx]):
        print("small case, algo %d Q %d hsize %d T %d" % (prm.algo, prm.Q, prm.hsize, prm.threads), flush=True)
        job = ctx.job(2, ks, cs, nmin, nmax, bm.copy(), wps, prm)
        job.sieve(p0, p1)
        ctx.sync()
        c = job.counters()
        got = job.bitmap()
        job.close()
        assert np.array_equal(got, ob)
        assert c.primecount == cnt.primecount
        if prm.algo == R.ALGO_THREE_LISTS:
            assert c.claims == cnt.claims


# Binary64 giant steps in whole workgroup rounds (F = M / T >= 1) below 2^51:
# n-spans of 3e5 and 6e5 give the reference split Q = 1024 and M = 293 or 586
# giant steps (parity lists: 147 or 293), so a 256-thread workgroup runs one or
# two full rounds and a 512-thread one a full round, plus a partial round; a
# one-entry queue makes every round boundary resolve (and redo) inside the
# full rounds.  The integer path gets the same shapes above 2^51.
FULL_ROUND_CASES = [
    # name, base, K, nmin, nmax, pmin, pmax, seed, density, threads, queue_cap
    ("f1_t256", 2, 8, 100000, 400000, 2**50, 2**50 + 2 * 10**5, 31, 0.5, 256, 0),
    ("f1_t256_q1", 2, 8, 100000, 400000, 2**50, 2**50 + 2 * 10**5, 31, 0.5, 256, 1),
    ("f2_t256", 2, 6, 100000, 700000, 2**44, 2**44 + 10**5, 32, 0.5, 256, 0),
    ("f1_t512", 2, 8, 100000, 700000, 2**50, 2**50 + 10**5, 33, 0.5, 512, 0),
    ("f1_t512_q1", 3, 6, 100000, 700000, 2**48, 2**48 + 10**5, 34, 0.5, 512, 1),
    ("f1_t128", 5, 4, 100000, 400000, 2**46, 2**46 + 10**5, 35, 0.5, 128, 0),
    ("int_f2_t256", 2, 6, 100000, 700000, 2**52, 2**52 + 10**5, 36, 0.5, 256, 0),
]


@pytest.mark.parametrize("case", FULL_ROUND_CASES, ids=[c[0] for c in FULL_ROUND_CASES])
def test_full_giant_rounds_match_oracle(ctx, case):
    name, base, K, nmin, nmax, pmin, pmax, seed, dens, threads, qcap = case
    print("full-round case", case, flush=True)
    ks, cs = _draw(base, K, seed)
    bm, wps = full_bitmap(K, nmin, nmax, density=dens, seed=seed)
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    M, QQ, MM = R.derive(prm, nmin, nmax)
    assert M >= threads  # at least one full round on the full list
    prm.threads = threads
    prm.queue_cap = qcap
    gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax, nthreads=8)
    _compare(base, ks, cs, nmin, nmax, wps, bm, gpu, orc)


@pytest.mark.parametrize("algo", [R.ALGO_MERGED, R.ALGO_CLASSES], ids=["merged", "classes"])
def test_order_path_with_more_targets_than_the_queue(ctx, algo):
    """Below p = 32768 the order of b is below a one-list table's span of
    32768 exponents, so every item takes the order path, which matches its
    targets against the baby values through the positive queue's LDS.  With
    100 sequences a one-list item holds up to 200 targets, more than the 128
    entries of that queue at 256 threads: they are matched in passes.  Base 5,
    seed 73 gives three primes in [3000, 30000) with 129 to 134 targets (the
    oracle's setup, counted below).  Bitmap and claims equal the oracle's."""
    base, K, nmin, nmax, pmin, pmax, seed = 5, 100, 20000, 120000, 3000, 30000, 73
    ks, cs = _draw(base, K, seed)
    job = O.make_job(base, ks, cs, nmin, nmax)
    most = 0
    for p in O.primes_in_range(pmin, pmax):
        pr, lst = O.setup_prime(job, int(p), 16384)
        if pr is not None:
            most = max(most, sum(2 if par == 1 else 1 for _, par, _ in lst))
    assert most > 128
    bm, wps = full_bitmap(K, nmin, nmax, density=0.5, seed=seed)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
    prm = merged_params(ctx, nmin, nmax, K, 16384, 16384, 256, algo=algo)
    gcnt, gfac, gbm = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    assert gcnt["order_lists"] > 0 and gcnt["seg_overflow"] == 0
    assert gcnt["claims"] == orc[0]["claims"]
    assert np.array_equal(gbm, orc[2])


def test_early_positives_over_the_staged_hash(ctx):
    """A filter of 64 words for tables of 64 entries passes a large share of
    the probes, so an item queues many positives within its first tiles: past
    slot 62 they would overwrite the W_j staged in the queue's top slots before
    every wavefront has hashed its share, and such an item, like one whose
    queue overflows, is redone exactly.  Bitmap and claims equal the oracle's
    (binary64 and integer paths)."""
    for base, K, nmin, nmax, pmin, pmax, seed in ((2, 20, 100000, 200000, 2**40, 2**40 + 10**5, 81),
                                                  (3, 12, 100000, 200000, 2**52, 2**52 + 4 * 10**4, 82)):
        print("early positives", base, K, pmin, flush=True)
        ks, cs = _draw(base, K, seed)
        bm, wps = full_bitmap(K, nmin, nmax, density=0.5, seed=seed)
        orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
        prm = merged_params(ctx, nmin, nmax, K, 64, 64, 256, algo=R.ALGO_CLASSES)
        gcnt, gfac, gbm = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
        assert gcnt["seg_overflow"] == 0
        assert gcnt["claims"] == orc[0]["claims"]
        assert np.array_equal(gbm, orc[2])


EDGE_CASES = [
    # name, base, K, nmin, nmax, pmin, pmax, seed, density
    ("empty_p_range", 2, 6, 20000, 60000, 10**8, 10**8, 21, 1.0),
    ("one_number_p_range", 2, 6, 20000, 60000, 10**8 + 7, 10**8 + 8, 22, 1.0),
    ("every_term_eliminated", 2, 10, 100000, 110000, 10**9, 10**9 + 2 * 10**6, 23, 0.0),
    ("single_n", 2, 8, 40000, 40000, 10**7, 10**7 + 10**6, 24, 1.0),
    ("two_n_base3", 3, 8, 40000, 40001, 10**7, 10**7 + 10**6, 25, 1.0),
    # the largest exponents a job takes (NMAX < 2^31) and bases up to 2^32 - 5
    ("n_below_2_31", 2, 6, 2**31 - 10002, 2**31 - 1, 10**7, 10**7 + 10**6, 27, 1.0),
    ("base_1000003", 1000003, 6, 20000, 40000, 10**8, 10**8 + 10**6, 28, 1.0),
    ("base_2_32_minus_5", 2**32 - 5, 6, 20000, 40000, 10**8, 10**8 + 10**6, 29, 1.0),
]


@pytest.mark.parametrize("case", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_edge_inputs_match_oracle(ctx, case):
    """Input edges: an empty p-range (no round runs), a p-range of one number,
    a bitmap whose terms are all eliminated already (every solution is a
    duplicate, nothing is claimed), n-spans of one and two exponents
    (solutions found past NMAX count as duplicates, giant.cl:21-39), the
    largest exponents and large bases.  The
    reference split is compared counter for counter; the one-list shapes on
    the cases with a full n-span."""
    name, base, K, nmin, nmax, pmin, pmax, seed, dens = case
    print("edge case", case, flush=True)
    ks, cs = _draw(base, K, seed)
    bm, wps = full_bitmap(K, nmin, nmax, density=dens, seed=seed)
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    gpu = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax, prm)
    orc = oracle_run(base, ks, cs, nmin, nmax, bm, pmin, pmax)
    _compare(base, ks, cs, nmin, nmax, wps, bm, gpu, orc)
    if pmax == pmin:
        assert gpu[0]["primecount"] == 0 and np.array_equal(gpu[2], bm)
    if dens == 0.0:
        assert gpu[0]["claims"] == 0 and gpu[0]["dups"] > 0 and not gpu[1].size
    if nmax - nmin >= 10000:
        for algo in (R.ALGO_MERGED, R.ALGO_CLASSES):
            print("edge case one-list", name, algo, flush=True)
            gcnt, gfac, gbm = _run_gpu(ctx, base, ks, cs, nmin, nmax, bm, wps, pmin, pmax,
                                       merged_params(ctx, nmin, nmax, K, algo=algo))
            assert gcnt["primecount"] == orc[0]["primecount"] and gcnt["claims"] == orc[0]["claims"]
            assert np.array_equal(gbm, orc[2])


def test_sieve_range_errors(ctx):
    """pmax below pmin and pmin below 3 are refused before anything runs."""
    ks, cs = _draw(2, 4, 26)
    bm, wps = full_bitmap(4, 20000, 30000)
    job = ctx.job(2, ks, cs, 20000, 30000, bm, wps, R.reference_params(20000, 30000, 4, 65536, ctx.compute_units))
    try:
        for lo, hi in ((10**6 + 1, 10**6), (2, 10**6), (0, 0)):
            with pytest.raises(R.RSError):
                job.sieve(lo, hi)
            with pytest.raises(R.RSError):
                job.sieve_noalloc(lo, hi)
        with pytest.raises(R.RSError):
            R.sieve_batch([job], [10**6 + 1], [10**6])
        assert job.counters().as_dict()["primecount"] == 0
        assert np.array_equal(job.bitmap(), bm)
    finally:
        job.close()
