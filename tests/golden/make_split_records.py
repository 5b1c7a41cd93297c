"""Writes tests/golden/split_records.npz: batch-split timings of synthetic
jobs (the training and held-out data of the cost model's split head).

Run on a GPU from the repository root:
    python tests/golden/make_split_records.py [--n 384] [--out tests/golden/split_records.npz]

Each record: the five features of a job descriptor drawn by
rssieve_amd.tuning.draw_descriptors (seed 8192) and the measured seconds per
prime of every (views per batched round, primes per segment) option of
rssieve_amd.tuning.SPLIT_OPTIONS (the library's rs_autotune_split on a
sample of about 2^22 primes at the job's p), with the job configured as the
committed candidates' model picks it.
"""
import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import rssieve_amd as R  # noqa: E402
from rssieve_amd import tuning  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=384)
    ap.add_argument("--seed", type=int, default=8192)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "split_records.npz"))
    ap.add_argument("--model", default=tuning.MODEL_PATH, help="the candidates' model that configures each job")
    a = ap.parse_args()
    ctx = R.Context(0)
    model = tuning.load_model(ctx, a.model)
    descs = tuning.draw_descriptors(a.n, a.seed)
    X, Y, V = tuning.collect_split(ctx, descs, model, log=sys.stdout)
    np.savez_compressed(a.out, X=X.astype(np.float32), Y=Y.astype(np.float32), V=V, descs=json.dumps(descs),
                        options=json.dumps(tuning.SPLIT_OPTIONS), device=ctx.name)
    print("wrote", a.out, X.shape, Y.shape, flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
