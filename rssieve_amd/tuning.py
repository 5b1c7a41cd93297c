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


def collect_split(ctx, descs, model, options=None, primes=1 << 22, log=None):
    """For each descriptor: a synthetic job on the candidates' model pick,
    timed by the library's rs_autotune_split on every batch-split option over
    a sample of about 'primes' primes at the job's p.  Returns (X, seconds per
    prime [n x options], valid mask)."""
    options = options or SPLIT_OPTIONS
    views = [v for v, _ in options]
    segp = [sp for _, sp in options]
    X, Y, V = [], [], []
    t0 = time.time()
    for i, d in enumerate(descs):
        nmax = d["nmin"] + d["span"] - 1
        prm = R.reference_params(d["nmin"], nmax, d["K"], 65536, ctx.compute_units)
        job, nterms, nmin, nmax = synth_job(ctx, d, prm)
        cands = ctx.tune_candidates(nmin, nmax, d["K"])
        x = features(d["base"], d["K"], nterms, d["span"], d["log2p"])
        if model is not None and model.n_out == len(cands):
            job.reconfigure(cands[int(model.choose(x[None, :], np.array([[c.threads != 0 for c in cands]]))[0])])
        pmin = int(2 ** d["log2p"])
        width = sample_width(d["log2p"], primes)
        best, secs, pr = job.autotune_split(pmin, width, views, segp)
        job.close()
        sp = np.array([sc / max(p_, 1) if sc > 0 else np.nan for sc, p_ in zip(secs, pr)], np.float64)
        X.append(x)
        Y.append(sp)
        V.append(np.isfinite(sp))
        if log and (i % 32 == 0 or i == len(descs) - 1):
            print("collect_split %d/%d  %.1fs  best=%s  sec/prime best %.3g one view %.3g" % (
                i + 1, len(descs), time.time() - t0, options[best], np.nanmin(sp), sp[0]), file=log, flush=True)
    return np.array(X), np.array(Y), np.array(V)


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


MODEL_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model")
MODEL_PATH = os.path.join(MODEL_DIR, "cost_model.bin")
META_PATH = os.path.join(MODEL_DIR, "cost_model.json")

# Batch-split options of the split head and of rs_autotune_split: (views per
# batched round, target primes per segment); one persistent BSGS launch
# serves views * seg_primes primes
SPLIT_OPTIONS = [(1, 65536), (4, 65536), (8, 65536), (16, 65536), (32, 65536), (64, 32768), (16, 131072),
                 (32, 131072)]


class SplitModel(CostModel):
    """The batch-split head: log(seconds per prime) of a job under one split
    option, relative to the job's mean over the options, from the five job
    features and the option itself (log2 views, log2 primes per segment): an
    MLP of 7 inputs and one output that every option shares, so that the
    trend with the launch size is learned across options rather than per
    option.  The pick is the option with the lowest prediction."""

    N_IN = N_FEATURES + 2

    def __init__(self, ctx, options=SPLIT_OPTIONS, hidden=8, seed=1):
        self.ctx = ctx
        self.options = [tuple(o) for o in options]
        self.mlp = ctx.mlp(self.N_IN, hidden, 1, seed=seed)
        self.mu = np.zeros(self.N_IN, np.float32)
        self.sd = np.ones(self.N_IN, np.float32)
        self.ymu = 0.0
        self.trained = False
        self.n_out = 1

    def _expand(self, X):
        X = np.asarray(X, np.float32)
        k = len(self.options)
        of = np.log2(np.array(self.options, np.float32))
        return np.concatenate([np.repeat(X, k, 0), np.tile(of, (len(X), 1))], 1).astype(np.float32)

    # 2000 epochs: the lowest 4-fold cross-validated regret on the 256
    # training jobs among widths 8, 16, 32 and 800-4000 epochs (0.20 %, 0.68 %
    # at 800 epochs)
    def fit(self, X, Y, V, epochs=2000, batch=256, lr=2e-3, seed=1):
        A = self._expand(X)
        self.mu = A.mean(0).astype(np.float32)
        self.sd = (A.std(0) + 1e-6).astype(np.float32)
        L = np.log(np.where(V, Y, 1.0))
        rel = L - (np.where(V, L, 0.0).sum(1, keepdims=True) / np.maximum(V.sum(1, keepdims=True), 1))
        T = np.where(V, rel, 0.0).reshape(-1, 1).astype(np.float32)
        M = V.reshape(-1, 1).astype(np.float32)
        self.ymu = 0.0
        loss = self.mlp.train(self._norm(A), T, M, epochs=epochs, batch=batch, lr=lr, seed=seed)
        self.trained = True
        return loss

    def predict(self, X):
        return self.mlp.forward(self._norm(self._expand(X))).reshape(len(X), len(self.options)) + self.ymu

    def choose(self, X, valid=None):
        P = self.predict(X)
        if valid is not None:
            P = np.where(valid, P, np.inf)
        return np.argmin(P, axis=1)

    def choose_split(self, X):
        """(views, seg_primes) of the predicted-fastest option per row."""
        return [self.options[i] for i in self.choose(X)]


def _head_bytes(kind, m, options=None):
    import struct
    params = np.concatenate([np.asarray(p, np.float32).reshape(-1) for p in m.mlp.get_params()])
    out = struct.pack("<4I", kind, m.mlp.n_in, m.mlp.H, m.mlp.n_out)
    out += np.asarray(m.mu, "<f4").tobytes() + np.asarray(m.sd, "<f4").tobytes()
    out += struct.pack("<dI", float(m.ymu), params.size) + params.astype("<f4").tobytes()
    if kind == 1:
        out += struct.pack("<I", len(options))
        out += np.array([v for v, _ in options], "<u4").tobytes() + np.array([s for _, s in options], "<u4").tobytes()
    return out


def save_model(m, path=MODEL_PATH, meta=None, split=None, meta_path=None):
    """The model file rs_model_load reads (include/rssieve_mlp.h): the
    candidates' head and, if given, the batch-split head; meta goes to a JSON
    file beside it."""
    import struct
    os.makedirs(os.path.dirname(path), exist_ok=True)
    heads = [_head_bytes(0, m)] + ([_head_bytes(1, split, split.options)] if split is not None else [])
    with open(path, "wb") as f:
        f.write(b"RSCM" + struct.pack("<2I", 1, len(heads)) + b"".join(heads))
    if meta is not None:
        with open(meta_path or os.path.splitext(path)[0] + ".json", "w") as f:
            json.dump(meta, f, indent=1, sort_keys=True)


def read_model_file(path=MODEL_PATH):
    """The heads of a model file as dicts of numpy arrays (no device work)."""
    import struct
    raw = open(path, "rb").read()
    assert raw[:4] == b"RSCM", "not a cost model file"
    version, nheads = struct.unpack_from("<2I", raw, 4)
    assert version == 1
    o, heads = 12, []
    for _ in range(nheads):
        kind, n_in, H, n_out = struct.unpack_from("<4I", raw, o)
        o += 16
        mu = np.frombuffer(raw, "<f4", n_in, o).copy()
        o += 4 * n_in
        sd = np.frombuffer(raw, "<f4", n_in, o).copy()
        o += 4 * n_in
        ymu, npar = struct.unpack_from("<dI", raw, o)
        o += 12
        flat = np.frombuffer(raw, "<f4", npar, o).copy()
        o += 4 * npar
        shapes = [(n_in, H), (H,), (H, H), (H,), (H, n_out), (n_out,)]
        params, k = [], 0
        for sh in shapes:
            n = int(np.prod(sh))
            params.append(flat[k:k + n].reshape(sh))
            k += n
        head = {"kind": kind, "n_in": n_in, "hidden": H, "n_out": n_out, "mu": mu, "sd": sd, "ymu": ymu,
                "params": params}
        if kind == 1:
            (nopt,) = struct.unpack_from("<I", raw, o)
            o += 4
            v = np.frombuffer(raw, "<u4", nopt, o)
            o += 4 * nopt
            sp = np.frombuffer(raw, "<u4", nopt, o)
            o += 4 * nopt
            head["options"] = list(zip(v.tolist(), sp.tolist()))
        heads.append(head)
    return heads


def _restore(m, head):
    m.mlp.set_params(head["params"])
    m.mu, m.sd, m.ymu = head["mu"].astype(np.float32), head["sd"].astype(np.float32), float(head["ymu"])
    m.trained = True
    return m


def load_model(ctx, path=MODEL_PATH):
    """The committed cost model's candidate head on this context (raises if
    there is none)."""
    head = next(h for h in read_model_file(path) if h["kind"] == 0)
    return _restore(CostModel(ctx, head["n_out"], hidden=head["hidden"]), head)


def load_split_model(ctx, path=MODEL_PATH):
    """The committed model's batch-split head, or None."""
    head = next((h for h in read_model_file(path) if h["kind"] == 1), None)
    if head is None:
        return None
    return _restore(SplitModel(ctx, head["options"], hidden=head["hidden"]), head)


class CModel:
    """The model file through the library's C ABI (rs_model_load,
    rs_model_choose, rs_model_split): the path the command line takes."""

    def __init__(self, ctx, path=MODEL_PATH):
        import ctypes as C
        self._C = C
        h = C.c_void_p()
        R._check(R.lib().rs_model_load(ctx._h, path.encode(), C.byref(h)))
        self._h = h

    def choose(self, base, kcount, nterms, span, log2p, cands, threads_only=0):
        C = self._C
        arr = (R.Params * len(cands))(*cands)
        idx = C.c_int()
        R._check(R.lib().rs_model_choose(self._h, base, kcount, nterms, span, log2p, arr, len(cands), threads_only,
                                         C.byref(idx)))
        return idx.value

    def rank(self, base, kcount, nterms, span, log2p, cands, k=3, threads_only=0):
        """The k eligible candidates with the lowest predictions, fastest first
        (rs_model_rank)."""
        C = self._C
        arr = (R.Params * len(cands))(*cands)
        idx = (C.c_int * k)()
        n = C.c_int()
        R._check(R.lib().rs_model_rank(self._h, base, kcount, nterms, span, log2p, arr, len(cands), threads_only, k,
                                       idx, C.byref(n)))
        return [idx[i] for i in range(n.value)]

    def split(self, base, kcount, nterms, span, log2p):
        C = self._C
        v, sp = C.c_uint32(), C.c_uint32()
        R._check(R.lib().rs_model_split(self._h, base, kcount, nterms, span, log2p, C.byref(v), C.byref(sp)))
        return v.value, sp.value

    def close(self):
        if self._h:
            R.lib().rs_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def model_pick(ctx, cands, base, kcount, nterms, span, log2p, path=MODEL_PATH):
    """The committed cost model's choice among the candidate list, or None
    when there is no model or it was trained on another candidate list."""
    try:
        model = load_model(ctx, path)
    except (OSError, StopIteration, AssertionError):
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
    t.add_argument("--rows", type=int, default=0, help="train on the first ROWS records only (0: all)")
    t.add_argument("--split-records", default=None, help="also train the batch-split head on these records")
    t.add_argument("--split-rows", type=int, default=0)
    a = ap.parse_args(argv)
    ctx = R.Context(0)
    if a.cmd == "collect":
        descs = draw_descriptors(a.n, a.seed)
        X, Y, V, cands = collect(ctx, descs, log=sys.stdout)
        np.savez_compressed(a.out, X=X, Y=Y, V=V, descs=json.dumps(descs),
                            cands=json.dumps([p.as_dict() for p in cands]))
        print("wrote", a.out, X.shape, Y.shape)
    elif a.cmd == "train":
        # the rows after --rows are the held-out rows of the tests: the
        # shipped model never sees them
        z = np.load(a.records, allow_pickle=False)
        n = a.rows or len(z["X"])
        X, Y, V = z["X"][:n], z["Y"][:n], z["V"][:n]
        m = CostModel(ctx, Y.shape[1], seed=a.seed)
        loss = m.fit(X, Y, V, seed=a.seed)
        meta = {"records": os.path.basename(a.records), "rows": int(len(X)), "candidates": int(Y.shape[1]),
                "final_loss": float(loss)}
        split = None
        if a.split_records:
            zs = np.load(a.split_records, allow_pickle=False)
            ns = a.split_rows or len(zs["X"])
            options = [tuple(o) for o in json.loads(str(zs["options"]))]
            split = SplitModel(ctx, options, seed=a.seed)
            sloss = split.fit(zs["X"][:ns], zs["Y"][:ns], zs["V"][:ns], seed=a.seed)
            meta.update({"split_records": os.path.basename(a.split_records), "split_rows": int(ns),
                         "split_options": options, "split_final_loss": float(sloss)})
        save_model(m, a.out, meta=meta, split=split)
        print(json.dumps(dict(meta, saved=a.out)))
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
