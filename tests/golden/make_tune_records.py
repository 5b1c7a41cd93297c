"""Writes tests/golden/tune_records.npz: autotuner timings of synthetic jobs
(the training and held-out data of the cost model).

Run on a GPU from the repository root:
    python tests/golden/make_tune_records.py [--n 4096] [--out tests/golden/tune_records.npz]
or in parts (--start/--count) merged afterwards with --merge part1.npz part2.npz.

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
    ap.add_argument("--n", type=int, default=4096, help="descriptors drawn")
    ap.add_argument("--seed", type=int, default=4096)
    ap.add_argument("--start", type=int, default=0, help="first descriptor of this run")
    ap.add_argument("--count", type=int, default=0, help="descriptors of this run (0: all from start)")
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden", "tune_records.npz"))
    ap.add_argument("--merge", nargs="*", default=None,
                    help="instead of collecting: concatenate these partial record files into --out")
    a = ap.parse_args()
    if a.merge:
        parts = [np.load(f, allow_pickle=False) for f in a.merge]
        descs = sum((json.loads(str(z["descs"])) for z in parts), [])
        np.savez_compressed(a.out, X=np.concatenate([z["X"] for z in parts]), Y=np.concatenate([z["Y"] for z in parts]),
                            V=np.concatenate([z["V"] for z in parts]), descs=json.dumps(descs),
                            cands=str(parts[0]["cands"]), device=str(parts[0]["device"]))
        print("merged", len(descs), "records into", a.out, flush=True)
        return
    ctx = R.Context(0)
    descs = tuning.draw_descriptors(a.n, a.seed)
    descs = descs[a.start:a.start + a.count] if a.count else descs[a.start:]
    X, Y, V, cands = tuning.collect(ctx, descs, log=sys.stdout)
    np.savez_compressed(a.out, X=X.astype(np.float32), Y=Y.astype(np.float32), V=V, descs=json.dumps(descs),
                        cands=json.dumps([p.as_dict() for p in cands]), device=ctx.name)
    print("wrote", a.out, X.shape, Y.shape, flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
