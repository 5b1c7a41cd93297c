#!/usr/bin/env python3
"""Benchmark: primes tested per second on one MI355X at the production shape.

Workload (one GPU): base 2, 100 sequences k*2^n+-1 (k odd, uniform in
[3, 2^32) from seed 2026, c = -1 for the first 50 and +1 for the rest),
n in [3*10^6, 10^7) with every term that has a prime factor < 2^20 removed
(pre-sieved on the CPU by the library's rs_presieve), sieved over
consecutive slices of p in [2^50, 2^50 + 2^38).  One step = one slice of
2^28 integers (~7.7 M primes); 'primes tested' is the reference's 'prime
count' (2-PRPs generated, setup.cl:99-101).  Inputs are resident in HBM before
the timed regions, which are bracketed by device synchronisation.

Two paths sieve the same slices of the same terms:
  plain  the reference's BSGS split (three tables, Q from build_giant_kernels,
         256-thread workgroups), one job, one round per segment;
  tuned  the configuration chosen for this job (the learned cost model when
         a trained model is committed, else the autotuner), the job split
         into 'batch' views sieved as one batch: one persistent BSGS launch
         per round over all of them, rounds replayed from a HIP graph.
Their final bitmaps must be identical (checked).  'value' is the tuned path.
A third leg ('jobs16') sieves sixteen uneven synthetic jobs one at a time
with the reference split and then all in one batch with the cost model's
picks, and compares the bitmaps.  A fourth ('high_p') sieves the high-p case
(b=5, K=100, p from 2^63: integer Montgomery products) with the reference
split and with the cost model's configuration and split, and reports that
kernel's multiply rate against the measured v_mad_u64_u32 peak.  Every factor
drained is checked with the exact 128-bit test.  With --other-workloads, the
small case (b=2, K=10, p from 10^9) and the high-p case are timed one job
each, reference split against the cost model's pick.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import rssieve_amd as R  # noqa: E402

NMIN, NMAX = 3000000, 9999999
P0 = 2**50
PRANGE = 2**38
SLICE = 2**28
K = 100
ALGO_NAMES = {0: "three lists", 1: "merged", 2: "residue classes"}
SEED = 2026
PRESIEVE = 2**20
# Every modular product of the BSGS kernel on this p-range (p < 2^51) is the
# binary64 sequence of rs_mont.h fp_mulmod: one mul (1 flop), three fma (2
# flops each), three adds (1 each) = 10 FP64 flops.  MI355X's published vector
# FP64 peak is 78.6 TFLOP/s (256 CUs x 64 lanes x 2 flops x 2.4 GHz).
FLOPS_PER_MODMUL = 10
FP64_PEAK = 78.6e12
# Above 2^51 a product is rs_mont.h's integer Montgomery mmul: 8 v_mad_u64_u32,
# 4 v_mul_lo_u32 and 2 v_mul_hi_u32 (the gfx950 code of the library), 14
# 32x32-bit multiply instructions, against the device's measured
# v_mad_u64_u32 issue rate (rs_peak_mad_u64_u32).
INT_MULS_PER_MODMUL = 14


def make_workload():
    rng = np.random.default_rng(SEED)
    ks = []
    while len(ks) < K:
        k = int(rng.integers(3, 2**32)) | 1
        if k not in ks:
            ks.append(k)
    cs = [-1] * (K // 2) + [1] * (K - K // 2)
    nmin = NMIN & ~1
    wps = (NMAX - nmin + 1 + 31) // 32
    bits = np.zeros(K * wps, np.uint32)
    full = np.full(wps, 0xFFFFFFFF, np.uint32)
    tail = (NMAX - nmin + 1) % 32
    if tail:
        full[-1] = (1 << tail) - 1
    for i in range(K):
        bits[i * wps:(i + 1) * wps] = full
    nthreads = max(1, min(16, os.cpu_count() or 1))
    R.presieve(2, ks, cs, nmin, NMAX, bits, wps, PRESIEVE, nthreads=nthreads)
    return ks, cs, nmin, NMAX, bits, wps


def cpu_baseline(ks, cs, nmin, nmax, bits, target_seconds=15.0):
    """The CPU oracle (a port of the reference algorithm, oracle/rs_oracle.c)
    threaded over primes on a bounded sample of the same workload.  The
    threads are the host CPU share of one GPU (16 on the MI355X boxes, whose
    os.cpu_count() shows the whole host), or all cores of a smaller host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nthreads = max(1, min(16, os.cpu_count() or 1))
    job = O.make_job(2, ks, cs, nmin, nmax)
    lo = P0 + PRANGE - 2**30  # a slice the GPU steps do not reach
    width = 2**18
    spent, primes, total_w = 0.0, 0, 0
    while spent < target_seconds and total_w < 2**30:
        b = bits.copy()
        t = time.perf_counter()
        cnt, _, _ = O.sieve(job, lo + total_w, lo + total_w + width, b, fac_cap=1 << 16, nthreads=nthreads)
        dt = time.perf_counter() - t
        spent += dt
        primes += cnt.primecount
        total_w += width
        rate = cnt.primecount / max(dt, 1e-9)
        remaining = target_seconds - spent
        width = int(min(2**28, max(2**18, remaining * rate * 34.7 * 0.8)))
    return {"value": primes / spent, "unit": "primes/s", "cores": nthreads, "kind": "port",
            "sample": "oracle on p in [2^50+2^38-2^30, +%d): %d primes, %.1f s; %d threads, the CPU share of one "
                      "GPU (os.cpu_count() %d)" % (total_w, primes, spent, nthreads, os.cpu_count() or 0)}


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes (hipMemcpy), whose FETCH_SIZE and
    WRITE_SIZE counters scale the BSGS kernel's to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_bsgs_cfg2_traffic.txt")


def measured_traffic(path, kernel, primes_per_launch):
    """HBM bytes per launch of the BSGS kernel on this path: the bytes per
    prime of the committed PMC summary (FETCH_SIZE + WRITE_SIZE per launch,
    scaled by a copy of known size, over the primes of that launch) times this
    run's primes per launch, or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_prime %s %s:" % (path, kernel)):
                return float(line.split(":")[1]) * primes_per_launch
    except OSError:
        return None
    return None


def roofline(work_mmul, bsgs_s, launches, kernel, path, primes):
    """The BSGS kernel against its compute roofline: algorithmic modular
    products (the kernel's own count: baby steps, giant-step products,
    giant-step advances) x 10 FP64 flops, per average launch duration (HIP
    events on the library's stream).  The products run on the vector FP64
    units (this path has no matrix work; MI355X's vector FP64 peak is 78.6
    TFLOP/s), so the bound is the VALU."""
    avg = bsgs_s / max(launches, 1)
    per_launch = work_mmul / max(launches, 1)
    achieved = per_launch * FLOPS_PER_MODMUL / max(avg, 1e-12) / 1e12
    return {"bound": "valu", "compute": "fp64 VALU", "achieved": achieved, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": achieved / (FP64_PEAK / 1e12),
            "traffic": measured_traffic(path, kernel, primes / max(launches, 1)), "kernel": kernel,
            "launches": launches,
            "avg_launch_ms": avg * 1e3, "modmul_per_launch": per_launch,
            "modmul_per_s": per_launch / max(avg, 1e-12)}


def roofline_int(work_mmul, bsgs_s, launches, kernel, peak_mad, path, primes):
    """The BSGS kernel on the integer Montgomery path (p >= 2^51): its
    32x32-bit multiply instructions (14 per algorithmic product) per second
    against the measured v_mad_u64_u32 issue rate of the device."""
    avg = bsgs_s / max(launches, 1)
    per_launch = work_mmul / max(launches, 1)
    achieved = per_launch * INT_MULS_PER_MODMUL / max(avg, 1e-12)
    return {"bound": "valu", "compute": "int32 multiply (v_mad_u64_u32 / v_mul_lo_u32 / v_mul_hi_u32)",
            "achieved": achieved / 1e12, "peak": peak_mad / 1e12, "unit": "Tmul/s", "frac": achieved / peak_mad,
            "traffic": measured_traffic(path, kernel, primes / max(launches, 1)), "kernel": kernel, "launches": launches,
            "avg_launch_ms": avg * 1e3,
            "modmul_per_launch": per_launch, "peak_source": "rs_peak_mad_u64_u32 on this device"}


def drain_verified(job, base, cap=1 << 20):
    """Drain the job's factors and check every one with the exact 128-bit test
    (rs_verify_factor, outside the timed regions): (factors, exact, composite
    2-PRPs); a false factor aborts the benchmark."""
    facs = job.drain(cap)
    ok = prp = 0
    for f in facs:
        k, c = job.sequence(int(f["kidx"]))
        v = R.verify_factor(int(f["p"]), k, int(f["n"]), c, base)
        if v == 0:
            sys.exit("bench.py: %d does not divide %d*%d^%d%+d" % (int(f["p"]), k, base, int(f["n"]), c))
        ok += v == 1
        prp += v == -1
    return len(facs), ok, prp


def run_plain(ctx, data, steps, warmup):
    ks, cs, nmin, nmax, bits, wps = data
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    prm.factor_cap = 1 << 20
    job = ctx.job(2, ks, cs, nmin, nmax, bits, wps, prm)
    p = P0
    for _ in range(warmup):
        job.sieve(p, p + SLICE)
        p += SLICE
    c0 = job.counters()
    job.timer_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        job.sieve(p, p + SLICE)
        p += SLICE
    ctx.sync()
    t1 = time.perf_counter()
    c1 = job.counters()
    bsgs_s, launches, _ = job.kernel_time()
    bm = job.bitmap()
    nf, ok, prp = drain_verified(job, 2)
    job.close()
    M, QQ, MM = R.derive(prm, nmin, nmax)
    return {"primes": c1.primecount - c0.primecount, "seconds": t1 - t0, "bitmap": bm,
            "roof": roofline(c1.work_mmul - c0.work_mmul, bsgs_s, launches, "k_bsgs<%d>" % prm.threads, "plain",
                             c1.primecount - c0.primecount),
            "busy": bsgs_s / (t1 - t0), "claims": c1.claims, "dups": c1.dups, "verified": (nf, ok, prp),
            "cfg": {"algo": "three lists", "Q": prm.Q, "M": M, "QQ": QQ, "MM": MM, "hsize": prm.hsize,
                    "threads": prm.threads, "seg_primes": prm.seg_primes}}


def model_choice(ctx, base, kcount, nterms, span, log2p, cands, threads_only=0, k=3):
    """The committed cost model's k fastest candidates (fastest first) and
    batch split through the C ABI (rs_model_rank, rs_model_split: the command
    line's path), or (None, None) without a model; a model without a split
    head gives split None."""
    from rssieve_amd import tuning
    try:
        cm = tuning.CModel(ctx)
    except R.RSError:
        return None, None
    try:
        idx = cm.rank(base, kcount, nterms, span, log2p, cands, k, threads_only)
    except R.RSError:
        cm.close()
        return None, None
    try:
        split = cm.split(base, kcount, nterms, span, log2p)
    except R.RSError:
        split = None
    cm.close()
    return idx, split


def measured_pick(job, cands, top, pmin, width, log):
    """The fastest of the model's top candidates, timed on the sample [pmin,
    pmin + width) of the job (rs_autotune: the job's bitmap and counters are
    restored), as the command line's -l does."""
    if len(top) == 1:
        return top[0]
    best, secs, primes = job.autotune(pmin, width, [cands[i] for i in top])
    log["model_top"] = list(top)
    log["model_top_ns_per_prime"] = [round(1e9 * s / max(n, 1), 3) for s, n in zip(secs, primes)]
    return top[best]


def choose_config(ctx, job, data, log):
    """The tuned configuration and batch split: the committed cost model's
    picks when a model is committed, else the autotuner's measured best on one
    sample slice and 32 views per round."""
    ks, cs, nmin, nmax, bits, wps = data
    cands = ctx.tune_candidates(nmin, nmax, K)
    nterms = int(np.bitwise_count(bits).sum())
    top, split = model_choice(ctx, 2, K, nterms, nmax - nmin + 1, np.log2(float(P0)), cands)
    if top is not None:
        how = "cost model top 3, timed"
        idx = measured_pick(job, cands, top, P0 + PRANGE - 2**31, 2**26, log)
        log["model_pick"] = idx
        log["model_split"] = split
    else:
        how = "autotune"
        best, secs, primes = job.autotune(P0 + PRANGE - 2**31, 2**26, cands)
        idx = best
        log["autotune_seconds"] = [round(x, 5) for x in secs]
    return cands[idx], how, idx, split


def _small_case():
    """The reference's CPU-sized case: b=2, ten odd k in [3, 10^6) (seed 1), c = -1 for five and +1
    for five, n in [10^5, 2*10^5) with the terms that have a prime factor below 2^16 removed."""
    rng = np.random.default_rng(1)
    ks = sorted(set(int(k) | 1 for k in rng.integers(3, 10**6, 10)))[:10]
    cs = [-1] * 5 + [1] * 5
    nmin, nmax = 100000, 199999
    wps = (nmax - nmin + 1 + 31) // 32
    bm = np.full(len(ks) * wps, 0xFFFFFFFF, np.uint32)
    tail = (nmax - nmin + 1) % 32
    if tail:
        bm.reshape(len(ks), wps)[:, -1] &= np.uint32((1 << tail) - 1)
    R.presieve(2, ks, cs, nmin, nmax, bm, wps, 2**16)
    return ks, cs, nmin, nmax, bm, wps


def _high_p_case(seed=63, K=100, nmin=10**6, nmax=2**24 - 1):
    """b=5, 100 sequences with k in [2, 2^32) coprime to 5 (seed 63), c alternating -1, +1,
    n in [10^6, 2^24), a bitmap keeping about 1/32 of the terms (the share a pre-sieve to 2^20
    leaves; the AND of five random words per word)."""
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < K:
        k = int(rng.integers(2, 2**32))
        if k % 5 and k not in ks:
            ks.append(k)
    cs = [1 if i % 2 else -1 for i in range(K)]
    wps = (nmax - nmin + 1 + 31) // 32
    bm = rng.integers(0, 2**32, size=K * wps, dtype=np.uint32)
    for _ in range(4):
        bm &= rng.integers(0, 2**32, size=K * wps, dtype=np.uint32)
    tail = (nmax - nmin + 1) % 32
    if tail:
        bm.reshape(K, wps)[:, -1] &= np.uint32((1 << tail) - 1)
    return ks, cs, nmin, nmax, bm, wps


def run_other_workloads(ctx):
    """The other single-GPU workloads, one job each (not batched): the reference split against the
    cost model's pick on the same terms: the small case over all of p in [10^9, 1.1*10^9), and the
    high-p case (integer Montgomery products) on p in [2^63, 2^63 + 2^28)."""
    from rssieve_amd import tuning
    model = tuning.load_model(ctx)
    res = []
    for name, base, data, p0, p1 in (
            ("b=2, K=10, n in [1e5,2e5) pre-sieved to 2^16, p in [1e9, 1.1e9)", 2, _small_case(), 10**9, 11 * 10**8),
            ("b=5, K=100, n in [1e6,2^24) (1/32 of the terms), p in [2^63, 2^63+2^28)", 5, _high_p_case(),
             2**63, 2**63 + 2**28)):
        ks, cs, nmin, nmax, bm, wps = data
        ref = R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units)
        cands = ctx.tune_candidates(nmin, nmax, len(ks))
        x = tuning.features(base, len(ks), int(np.bitwise_count(bm).sum()), nmax - nmin + 1, np.log2(float(p0)))
        idx = int(model.choose(x[None, :], np.array([[c.threads != 0 for c in cands]]))[0])
        out = {}
        for label, prm in (("plain", ref), ("tuned", cands[idx])):
            w = ctx.job(base, ks, cs, nmin, nmax, bm.copy(), wps, prm)  # warm-up of the same shape
            w.sieve(p0, p0 + (p1 - p0) // 64)
            ctx.sync()
            w.close()
            job = ctx.job(base, ks, cs, nmin, nmax, bm.copy(), wps, prm)
            t0 = time.perf_counter()
            job.sieve(p0, p1)
            ctx.sync()
            dt = time.perf_counter() - t0
            out[label] = (job.counters().primecount, dt, job.bitmap())
            job.close()
        res.append({"workload": name, "primes": int(out["plain"][0]),
                    "plain_primes_per_s": out["plain"][0] / out["plain"][1],
                    "tuned_primes_per_s": out["tuned"][0] / out["tuned"][1],
                    "tuned_over_plain": out["plain"][1] / out["tuned"][1],
                    "same_terms": bool(np.array_equal(out["plain"][2], out["tuned"][2])),
                    "candidate": idx, "tuned_algo": int(cands[idx].algo), "tuned_Q": int(cands[idx].Q)})
    return res


def run_jobs16(ctx, seed=16, width=2**26):
    """Sixteen uneven jobs (bases 2, 3, 5, 6, 10; 1-100 sequences and n-spans
    10^4..10^7 log-uniform; k odd below 2^31 coprime to the base; bitmaps of
    1-10 % density; p from 2^48, one slice of 2^26 each): one job at a time
    with the reference split, against all sixteen in one batch with the cost
    model's 256-thread pick per job.  Final bitmaps must agree."""
    from rssieve_amd import tuning
    descs = tuning.draw_descriptors(16, seed, log2p_range=(48.0, 48.0))
    try:
        model = tuning.load_model(ctx)
    except Exception:
        model = None

    cache = {}

    def make_jobs(tuned):
        jobs = []
        for d in descs:
            prm = R.reference_params(d["nmin"], d["nmin"] + d["span"] - 1, d["K"], 65536, ctx.compute_units)
            job, nterms, nmin, nmax = tuning.synth_job(ctx, d, prm, cache)
            if tuned and model is not None:
                cands = ctx.tune_candidates(nmin, nmax, d["K"])
                if len(cands) == model.n_out:
                    x = tuning.features(d["base"], d["K"], nterms, d["span"], 48.0)[None, :]
                    valid = np.array([[c.threads == 256 for c in cands]])
                    job.reconfigure(cands[int(model.choose(x, valid)[0])])
            jobs.append(job)
        return jobs

    p0 = 2**48
    out = {}
    for mode in ("sequential", "batched"):
        jobs = make_jobs(mode == "batched")
        ctx.sync()
        t0 = time.perf_counter()
        if mode == "sequential":
            for j in jobs:
                j.sieve(p0, p0 + width)
        else:
            R.sieve_batch(jobs, [p0] * len(jobs), [p0 + width] * len(jobs))
        ctx.sync()
        dt = time.perf_counter() - t0
        primes = sum(j.counters().primecount for j in jobs)
        out[mode] = {"primes_per_s": primes / dt, "seconds": dt, "bitmaps": [j.bitmap() for j in jobs]}
        for j in jobs:
            j.close()
    same = all(np.array_equal(a, b) for a, b in zip(out["sequential"]["bitmaps"], out["batched"]["bitmaps"]))
    return {"workload": "16 uneven jobs (seed %d), p in [2^48, 2^48 + 2^26) each" % seed,
            "sequential_reference_primes_per_s": out["sequential"]["primes_per_s"],
            "batched_tuned_primes_per_s": out["batched"]["primes_per_s"],
            "batched_over_sequential": out["batched"]["primes_per_s"] / out["sequential"]["primes_per_s"],
            "same_terms": bool(same), "chosen_by": "cost model" if model is not None else "reference split"}


def run_tuned(ctx, data, steps, warmup, batch):
    ks, cs, nmin, nmax, bits, wps = data
    prm = R.reference_params(nmin, nmax, K, 65536, ctx.compute_units)
    prm.factor_cap = 1 << 20
    job = ctx.job(2, ks, cs, nmin, nmax, bits, wps, prm)
    log = {}
    cfg, how, idx, split = choose_config(ctx, job, data, log)
    cfg.factor_cap = 1 << 20
    split_by = "option"
    if not batch:
        if split is not None:
            batch, cfg.seg_primes = split
            split_by = "cost model"
        else:
            batch = 32
    job.reconfigure(cfg)
    views = [job.view() for _ in range(batch - 1)]
    handles = [job] + views
    part = SLICE // batch

    def step(p):
        R.sieve_batch(handles, [p + i * part for i in range(batch)],
                      [p + (i + 1) * part if i < batch - 1 else p + SLICE for i in range(batch)])

    p = P0
    for _ in range(warmup):
        step(p)
        p += SLICE
    c0 = job.counters()
    job.timer_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(p)
        p += SLICE
    ctx.sync()
    t1 = time.perf_counter()
    c1 = job.counters()
    bsgs_s, launches, _ = job.kernel_time()
    bm = job.bitmap()
    nf, ok, prp = drain_verified(job, 2)
    for v in views:
        v.close()
    job.close()
    return {"primes": c1.primecount - c0.primecount, "seconds": t1 - t0, "bitmap": bm,
            "roof": roofline(c1.work_mmul - c0.work_mmul, bsgs_s, launches, "k_bsgs<%d>" % cfg.threads, "tuned",
                             c1.primecount - c0.primecount),
            "busy": bsgs_s / (t1 - t0), "claims": c1.claims, "dups": c1.dups, "verified": (nf, ok, prp),
            "cfg": {"algo": ALGO_NAMES[cfg.algo], "Q": cfg.Q,
                    "hsize": cfg.hsize, "threads": cfg.threads, "seg_primes": cfg.seg_primes, "batch": batch,
                    "chosen_by": how, "split_by": split_by, "candidate": idx},
            "log": log}


def run_high_p(ctx, peak_mad, width=2**28):
    """The high-p case on the integer Montgomery path: b=5, K=100, n in
    [10^6, 2^24) (1/32 of the terms), p in [2^63, 2^63 + width): the reference
    split (one job, one view) and the cost model's configuration and split
    (batched views), same final bitmap; the tuned kernel's roofline is its
    32x32-bit multiply rate against the measured v_mad_u64_u32 peak."""
    ks, cs, nmin, nmax, bm, wps = _high_p_case()
    p0 = 2**63
    base = 5
    nterms = int(np.bitwise_count(bm).sum())
    cands = ctx.tune_candidates(nmin, nmax, len(ks))
    top, split = model_choice(ctx, base, len(ks), nterms, nmax - nmin + 1, 63.0, cands)
    ref = R.reference_params(nmin, nmax, len(ks), 65536, ctx.compute_units)
    idx, tlog = None, {}
    if top is not None:
        probe = ctx.job(base, ks, cs, nmin, nmax, bm.copy(), wps, ref)
        idx = measured_pick(probe, cands, top, p0 + 2 * width, 2**25, tlog)  # a sample past the timed range
        probe.close()
    out = {}
    for label, prm, nv in (("plain", ref, 1), ("tuned", cands[idx if idx is not None else 0], 32)):
        prm.factor_cap = 1 << 16
        if label == "tuned" and split is not None:
            nv, prm.seg_primes = split
        job = ctx.job(base, ks, cs, nmin, nmax, bm.copy(), wps, prm)
        views = [job.view() for _ in range(nv - 1)]
        hs = [job] + views
        part = width // nv
        lo = [p0 + i * part for i in range(nv)]
        hi = [p0 + (i + 1) * part if i < nv - 1 else p0 + width for i in range(nv)]
        warm = [x + 1 for x in lo]  # the batch shape's graphs, outside the timed region
        R.sieve_batch(hs, lo, warm)
        ctx.sync()
        c0 = job.counters()
        job.timer_reset()
        t0 = time.perf_counter()
        R.sieve_batch(hs, [x + 1 for x in lo], hi)
        ctx.sync()
        dt = time.perf_counter() - t0
        c1 = job.counters()
        bsgs_s, launches, _ = job.kernel_time()
        nf, ok, prp = drain_verified(job, base, 1 << 16)
        out[label] = {"primes": c1.primecount - c0.primecount, "seconds": dt, "bitmap": job.bitmap(),
                      "roof": roofline_int(c1.work_mmul - c0.work_mmul, bsgs_s, launches, "k_bsgs<%d>" % prm.threads,
                                           peak_mad, "high_p_" + label, c1.primecount - c0.primecount),
                      "busy": bsgs_s / dt, "verified": ok, "algo": ALGO_NAMES[prm.algo], "Q": prm.Q,
                      "threads": prm.threads, "views": nv, "seg_primes": prm.seg_primes}
        for v in views:
            v.close()
        job.close()
    same = bool(np.array_equal(out["plain"]["bitmap"], out["tuned"]["bitmap"]))
    pl, tu = out["plain"], out["tuned"]
    return {"workload": "b=5, K=100, n in [1e6,2^24) (1/32 of the terms), p in [2^63, 2^63+2^28)",
            "plain_primes_per_s": pl["primes"] / pl["seconds"], "tuned_primes_per_s": tu["primes"] / tu["seconds"],
            "tuned_over_plain": pl["seconds"] / tu["seconds"] * tu["primes"] / max(pl["primes"], 1),
            "same_terms": same, "tuned_config": {k: tu[k] for k in ("algo", "Q", "threads", "views", "seg_primes")},
            "candidate": idx, "chosen_by": "cost model top 3, timed" if top is not None else "reference split",
            "model_top": tlog.get("model_top"), "model_top_ns_per_prime": tlog.get("model_top_ns_per_prime"),
            "factors_verified": pl["verified"], "primes": tu["primes"],
            "roofline": dict(tu["roof"], bsgs_busy_frac=tu["busy"]),
            "plain_roofline": dict(pl["roof"], bsgs_busy_frac=pl["busy"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=0,
                    help="views of the job per batched round (default: the cost model's split, else 32)")
    ap.add_argument("--no-high-p", action="store_true", help="skip the high-p (integer Montgomery) leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--tuned-only", action="store_true", help="skip the plain leg (for profiling the tuned kernel)")
    ap.add_argument("--no-jobs16", action="store_true", help="skip the 16-job batch leg")
    ap.add_argument("--other-workloads", action="store_true",
                    help="also time the small case and the high-p case (plain vs cost-model pick)")
    ap.add_argument("--high-p-only", action="store_true",
                    help="run only the high-p leg and print its record (for profiling its kernels)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first copy this many MiB device to device (a kernel of known traffic for PMC scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (several GPUs would each sieve their own "
              "p-range and merge the cleared terms at the end)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.steps + args.warmup > (PRANGE - 2**31) // SLICE:
        sys.exit("bench.py: too many steps for the p-range")

    t_setup = time.perf_counter()
    data = make_workload()
    nterms = int(np.bitwise_count(data[4]).sum())
    setup_s = time.perf_counter() - t_setup

    ctx = R.Context(0)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib << 20)
    if args.high_p_only:
        high_p = run_high_p(ctx, ctx.peak_mad_u64_u32())
        ctx.close()
        print(json.dumps({"high_p": high_p}))
        return
    plain = None if args.tuned_only else run_plain(ctx, data, args.steps, args.warmup)
    res = plain
    tuned = None
    if not args.plain_only:
        tuned = run_tuned(ctx, data, args.steps, args.warmup, args.batch)
        if plain is not None and not np.array_equal(tuned["bitmap"], plain["bitmap"]):
            sys.exit("bench.py: tuned and plain paths eliminated different terms")
        res = tuned
    if plain is None:
        plain = tuned
    plain_value = plain["primes"] / plain["seconds"]
    jobs16 = None
    if not args.no_jobs16 and not args.plain_only and not args.tuned_only:
        jobs16 = run_jobs16(ctx)
    other = run_other_workloads(ctx) if args.other_workloads else None
    high_p = None
    if not args.no_high_p and not args.plain_only and not args.tuned_only:
        peak_mad = ctx.peak_mad_u64_u32()
        high_p = run_high_p(ctx, peak_mad)
        if not high_p["same_terms"]:
            sys.exit("bench.py: high-p tuned and plain paths eliminated different terms")
    ctx.close()
    value = res["primes"] / res["seconds"]
    line = {
        "metric": "primes tested/sec, 100 seqs p~2^50: batched+learned-tuned vs plain GPU, same terms",
        "value": value,
        "unit": "primes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "b=2, K=100 odd k<2^32 (seed 2026), n in [3e6,1e7) pre-sieved to 2^20 "
                               "(%d terms), p slices of 2^28 from 2^50" % nterms,
                   "path": "tuned+batched" if tuned else "plain", **res["cfg"]},
        "roofline": dict(res["roof"], bsgs_busy_frac=res["busy"]),
        "plain_value": plain_value,
        "plain_config": plain["cfg"],
        "plain_roofline": dict(plain["roof"], bsgs_busy_frac=plain["busy"]),
        "tuned_over_plain": value / plain_value,
        "same_terms": tuned is not None and not args.tuned_only,
        "primes_per_step": res["primes"] / args.steps,
        "factors": int(res["claims"]),
        "factors_verified": int(res["verified"][1]),
        "setup_s": setup_s,
    }
    if tuned and not args.tuned_only:
        # The reference algorithm's work on the same primes: the plain leg
        # runs the reference split (three tables, giant.cl's loop bounds) and
        # counts its modular products exactly (the oracle's count, checked in
        # the parity tests).  At the tuned leg's rate that work would need
        # this much FP64 throughput: the effect of the residue-class tables,
        # shown beside the executed-work roofline above.
        ref_pp = plain["roof"]["modmul_per_launch"] * plain["roof"]["launches"] / max(plain["primes"], 1)
        ach = ref_pp * value * FLOPS_PER_MODMUL / 1e12
        line["roofline"]["reference_equivalent"] = {
            "modmul_per_prime_reference": ref_pp,
            "modmul_per_prime_executed": res["roof"]["modmul_per_launch"] * res["roof"]["launches"] / max(res["primes"], 1),
            "achieved": ach, "unit": "TFLOP/s", "frac": ach / (FP64_PEAK / 1e12)}
    if tuned:
        line["tuned_log"] = tuned["log"]
    if jobs16 is not None:
        line["jobs16"] = jobs16
    if other is not None:
        line["other_workloads"] = other
    if high_p is not None:
        line["high_p"] = high_p
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(*data[:5])
    print(json.dumps(line))


if __name__ == "__main__":
    main()
