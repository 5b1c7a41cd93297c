"""Configuration search for sieve jobs: the autotuner's timing records and the
learned cost model built on them.

  descriptor      (base, #sequences, #terms, n-span, log2 p) of a job
  collect()       for each descriptor, a synthetic job of that shape is timed
                  by the library's autotuner (rs_autotune) on every candidate
                  configuration (rs_tune_candidates): one record per job with
                  the seconds per prime of each candidate
  CostModel       the MLP of rssieve_mlp.h (5 -> 64 -> 64 -> #candidates,
                  fp32, trained and run by the library's HIP kernels) fitted to
                  log(seconds per prime); it picks the candidate with the
                  lowest prediction, and falls back to the reference split
                  (candidate 0) when it has no model
  regret()        time of the chosen candidate over the best measured one

Command line (on a GPU):
  python -m rssieve_amd.tuning collect --n 1024 --out records.npz
  python -m rssieve_amd.tuning evaluate --records records.npz
  python -m rssieve_amd.tuning train --records records.npz   (writes model/cost_model.npz)
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

import rssieve_amd as R

BASES = (2, 3, 5, 6, 10)
N_FEATURES = 5


def features(base, kcount, nterms, span, log2p):
    """The five model inputs (all logarithmic)."""
    return np.array([math.log2(base), math.log2(kcount), math.log2(nterms + 1.0), math.log2(span), log2p],
                    np.float32)


def draw_descriptors(n, seed, log2p_range=(32.0, 62.0)):
    """Job shapes: base from {2,3,5,6,10}, 1-100 sequences and n-spans
    10^4..10^7 log-uniform, term density 1-10 % (what a pre-sieve leaves),
    p from 2^32..2^62 log-uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        base = int(rng.choice(BASES))
        K = int(round(math.exp(rng.uniform(0.0, math.log(100.0)))))
        K = max(1, min(100, K))
        span = int(round(math.exp(rng.uniform(math.log(1e4), math.log(1e7)))))
        density = float(math.exp(rng.uniform(math.log(0.01), math.log(0.10))))
        log2p = float(rng.uniform(*log2p_range))
        nmin = int(rng.integers(1000, 10**7)) & ~1
        out.append({"base": base, "K": K, "span": span, "density": density, "log2p": log2p, "nmin": nmin,
                    "seed": int(rng.integers(0, 2**31))})
    return out


def synth_sequences(d, rng=None):
    """The sequences of descriptor d: K seeded k below 2^31 coprime to the
    base (odd for even bases, even for odd ones), c alternating -1, +1."""
    rng = np.random.default_rng(d["seed"]) if rng is None else rng
    base, K = d["base"], d["K"]
    ks = []
    while len(ks) < K:
        k = int(rng.integers(3, 2**31))
        if (k % 2 == 0) == (base % 2 == 0) or math.gcd(k, base) != 1 or k in ks:
            continue
        ks.append(k)
    cs = [1 if i % 2 else -1 for i in range(K)]
    return ks, cs


def synth_job(ctx, d, params, cache=None):
    """A job of descriptor d: the sequences of synth_sequences and a random
    bitmap whose bits are set with probability 2^-round(-log2 density).  A
    dict passed as cache keeps the generated bitmaps for later calls with the
    same descriptor."""
    rng = np.random.default_rng(d["seed"])
    base, K = d["base"], d["K"]
    ks, cs = synth_sequences(d, rng)
    nmin = d["nmin"]
    nmax = nmin + d["span"] - 1
    wps = (nmax - nmin + 1 + 31) // 32
    # every bit set with probability 2^-r ~ density: the AND of r random words
    key = (d["seed"], K, nmin, nmax)
    if cache is not None and key in cache:
        bits, nterms = cache[key]
        return ctx.job(base, ks, cs, nmin, nmax, bits, wps, params), nterms, nmin, nmax
    nwords = K * wps
    r = max(1, int(round(-math.log2(d["density"]))))
    bits = rng.integers(0, 2**32, size=nwords, dtype=np.uint32)
    for _ in range(r - 1):
        bits &= rng.integers(0, 2**32, size=nwords, dtype=np.uint32)
    tail = (nmax - nmin + 1) % 32
    if tail:
        bits.reshape(K, wps)[:, -1] &= np.uint32((1 << tail) - 1)
    nterms = int(np.bitwise_count(bits).sum())
    if cache is not None:
        cache[key] = (bits, nterms)
    return ctx.job(base, ks, cs, nmin, nmax, bits, wps, params), nterms, nmin, nmax


def sample_width(log2p, primes=65536):
    """p-range holding about 'primes' primes at 2^log2p."""
    return int(primes * log2p * math.log(2.0))


def collect(ctx, descs, log=None):
    """Time every candidate for every descriptor; returns (X, seconds per
    prime [n x C], valid mask, candidate list of the first job)."""
    X, Y, V = [], [], []
    cand0 = None
    t0 = time.time()
    for i, d in enumerate(descs):
        prm = R.reference_params(d["nmin"], d["nmin"] + d["span"] - 1, d["K"], 65536, ctx.compute_units)
        job, nterms, nmin, nmax = synth_job(ctx, d, prm)
        cands = ctx.tune_candidates(nmin, nmax, d["K"])
        if cand0 is None:
            cand0 = cands
        pmin = int(2 ** d["log2p"])
        width = sample_width(d["log2p"])
        best, secs, primes = job.autotune(pmin, width, cands)
        job.close()
        sp = np.array([s / max(p, 1) if s > 0 else np.nan for s, p in zip(secs, primes)], np.float64)
        X.append(features(d["base"], d["K"], nterms, d["span"], d["log2p"]))
        Y.append(sp)
        V.append(np.isfinite(sp))
        if log and (i % 32 == 0 or i == len(descs) - 1):
            print("collect %d/%d  %.1fs  best=%d  sec/prime best %.3g ref %.3g" % (
                i + 1, len(descs), time.time() - t0, best, np.nanmin(sp), sp[0]), file=log, flush=True)
    return np.array(X), np.array(Y), np.array(V), cand0


class CostModel:
    """log(seconds per prime) of every candidate from the five features."""

    def __init__(self, ctx, n_out, hidden=64, seed=1):
        self.ctx = ctx
        self.mlp = ctx.mlp(N_FEATURES, hidden, n_out, seed=seed)
        self.mu = np.zeros(N_FEATURES, np.float32)
        self.sd = np.ones(N_FEATURES, np.float32)
        self.ymu = 0.0
        self.trained = False
        self.n_out = n_out

    def _norm(self, X):
        return ((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)

    def fit(self, X, Y, V, epochs=400, batch=128, lr=2e-3, seed=1):
        self.mu = X.mean(0).astype(np.float32)
        self.sd = (X.std(0) + 1e-6).astype(np.float32)
        T = np.log(np.where(V, Y, 1.0))
        self.ymu = float(T[V].mean())
        T = np.where(V, T - self.ymu, 0.0).astype(np.float32)
        loss = self.mlp.train(self._norm(X), T, V.astype(np.float32), epochs=epochs, batch=batch, lr=lr, seed=seed)
        self.trained = True
        return loss

    def predict(self, X):
        return self.mlp.forward(self._norm(X)) + self.ymu

    def choose(self, X, valid=None):
        """Index of the predicted-fastest candidate per row (0, the reference
        split, when the model is untrained)."""
        if not self.trained:
            return np.zeros(len(X), np.int64)
        P = self.predict(X)
        if valid is not None:
            P = np.where(valid, P, np.inf)
        return np.argmin(P, axis=1)


MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model", "cost_model.npz")


def save_model(m, path=MODEL_PATH, meta=None):
    """Weights, input normalisation and output offset of a trained model."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    params = m.mlp.get_params()
    np.savez(path, mu=m.mu, sd=m.sd, ymu=np.float64(m.ymu), hidden=np.int64(m.mlp.H),
             n_out=np.int64(m.mlp.n_out), meta=json.dumps(meta or {}),
             **{"p%d" % i: p for i, p in enumerate(params)})


def load_model(ctx, path=MODEL_PATH):
    """The committed cost model on this context (raises if there is none)."""
    z = np.load(path, allow_pickle=False)
    m = CostModel(ctx, int(z["n_out"]), hidden=int(z["hidden"]))
    m.mlp.set_params([z["p%d" % i] for i in range(6)])
    m.mu, m.sd, m.ymu = z["mu"].astype(np.float32), z["sd"].astype(np.float32), float(z["ymu"])
    m.trained = True
    m.n_out = int(z["n_out"])
    return m


def model_pick(ctx, cands, base, kcount, nterms, span, log2p, path=MODEL_PATH):
    """The committed cost model's choice among the candidate list, or None
    when there is no model or it was trained on another candidate list."""
    try:
        model = load_model(ctx, path)
    except (OSError, KeyError):
        return None
    if model.n_out != len(cands):
        model.mlp.close()
        return None
    x = features(base, kcount, nterms, span, log2p)[None, :]
    idx = int(model.choose(x, np.array([[c.threads != 0 for c in cands]]))[0])
    model.mlp.close()
    return idx


def best_fixed(Y, V):
    """The non-learned rule fitted to the same records: the one candidate with
    the lowest mean log time over the rows where it ran."""
    L = np.where(V, np.log(np.where(V, Y, 1.0)), np.nan)
    score = np.nanmean(L, axis=0)
    score = np.where(V.mean(0) > 0.99, score, np.inf)  # valid for (almost) every job
    return int(np.argmin(score))


def regret(Y, V, choice):
    """time(chosen) / time(best measured) - 1, per row."""
    Yv = np.where(V, Y, np.inf)
    best = Yv.min(1)
    got = Yv[np.arange(len(Y)), choice]
    return got / best - 1.0


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--n", type=int, default=1024)
    c.add_argument("--seed", type=int, default=4096)
    c.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--records", required=True)
    e.add_argument("--test", type=int, default=256)
    e.add_argument("--out", default=None)
    t = sub.add_parser("train")
    t.add_argument("--records", required=True)
    t.add_argument("--out", default=MODEL_PATH)
    t.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    ctx = R.Context(0)
    if a.cmd == "collect":
        descs = draw_descriptors(a.n, a.seed)
        X, Y, V, cands = collect(ctx, descs, log=sys.stdout)
        np.savez_compressed(a.out, X=X, Y=Y, V=V, descs=json.dumps(descs),
                            cands=json.dumps([p.as_dict() for p in cands]))
        print("wrote", a.out, X.shape, Y.shape)
    elif a.cmd == "train":
        z = np.load(a.records, allow_pickle=False)
        X, Y, V = z["X"], z["Y"], z["V"]
        m = CostModel(ctx, Y.shape[1], seed=a.seed)
        loss = m.fit(X, Y, V, seed=a.seed)
        save_model(m, a.out, meta={"records": os.path.basename(a.records), "rows": int(len(X)),
                                   "candidates": int(Y.shape[1]), "final_loss": float(loss)})
        print(json.dumps({"saved": a.out, "rows": int(len(X)), "final_loss": float(loss)}))
    else:
        z = np.load(a.records, allow_pickle=False)
        X, Y, V = z["X"], z["Y"], z["V"]
        ntr = len(X) - a.test
        m = CostModel(ctx, Y.shape[1])
        loss = m.fit(X[:ntr], Y[:ntr], V[:ntr])
        ch = m.choose(X[ntr:], V[ntr:])
        r_model = regret(Y[ntr:], V[ntr:], ch)
        r_ref = regret(Y[ntr:], V[ntr:], np.zeros(len(ch), np.int64))
        fixed = best_fixed(Y[:ntr], V[:ntr])
        r_fix = regret(Y[ntr:], V[ntr:], np.full(len(ch), fixed, np.int64))
        res = {"fixed_candidate": fixed, "regret_fixed_mean": float(r_fix.mean()),
               "regret_fixed_median": float(np.median(r_fix)),"train_rows": int(ntr), "test_rows": int(len(ch)), "final_loss": float(loss),
               "regret_model_mean": float(r_model.mean()), "regret_model_median": float(np.median(r_model)),
               "regret_reference_mean": float(r_ref.mean()), "regret_reference_median": float(np.median(r_ref)),
               "speedup_model_vs_reference_geomean": float(np.exp(np.mean(np.log((1 + r_ref) / (1 + r_model)))))}
        print(json.dumps(res))
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f)
    ctx.close()


if __name__ == "__main__":
    main()
