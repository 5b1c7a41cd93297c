"""Writes tests/golden/tune_records.npz: autotuner timings of synthetic jobs
(the training and held-out data of the cost model).

Run on a GPU from the repository root:
    python tests/golden/make_tune_records.py [--n 1024] [--out tests/golden/tune_records.npz]

Each record: the five features of a job descriptor drawn by
rssieve_amd.tuning.draw_descriptors (seed 4096) and the measured seconds per
prime of every candidate configuration (rssieve_amd.tuning.collect, i.e. the
library's rs_autotune on a sample of ~65536 primes at the job's p).
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
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=4096)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "tune_records.npz"))
    a = ap.parse_args()
    ctx = R.Context(0)
    descs = tuning.draw_descriptors(a.n, a.seed)
    X, Y, V, cands = tuning.collect(ctx, descs, log=sys.stdout)
    np.savez_compressed(a.out, X=X.astype(np.float32), Y=Y, V=V, descs=json.dumps(descs),
                        cands=json.dumps([p.as_dict() for p in cands]), device=ctx.name)
    print("wrote", a.out, X.shape, Y.shape, flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
