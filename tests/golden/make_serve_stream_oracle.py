"""Writes tests/golden/serve_stream_oracle.npz: the oracle's results for a
sample of the integrals of a seeded serving stream (g3lhalo_amd.workloads.
serve_stream(8, seed=STREAM_SEED): prior HOD sets, 20 % two-population
requests at the three type pairs, 1-200 triplets each, random NNMap/NMapMap
term masks; tol 1e-4, maxEval 5e6): SAMPLE integrals per request, chosen
by a seeded draw over the request's integrals in submission order.  About a
minute on 8 cores.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_serve_stream_oracle.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from g3lhalo_amd import tables as T  # noqa: E402
from g3lhalo_amd import workloads as W  # noqa: E402

STREAM_SEED = 2024
N_REQUESTS = 8
SAMPLE = 3


def sample_positions(n_integrals: int, request: int) -> np.ndarray:
    rng = np.random.default_rng(STREAM_SEED * 1000 + request)
    return np.sort(rng.choice(n_integrals, size=min(SAMPLE, n_integrals), replace=False))


def main():
    O.build()
    t = T.make_tables()
    stream = W.serve_stream(N_REQUESTS, seed=STREAM_SEED)
    rows = []
    t0 = time.time()
    for i, q in enumerate(stream):
        m = O.Model(t)
        m.set_hods(q["hod1"], q["hod2"], identical=not q["two"])
        ints = W.request_integrals(q)
        for pos in sample_positions(len(ints), i):
            stat, term, th, t1, t2 = ints[pos]
            r = m.eval_term(stat, term, th * T.ARCMIN_IN_RAD, t1, t2, 1e-4, 5000000)
            rows.append((i, pos, stat, term, r["value"], r["raw"], r["err"], r["neval"],
                         r["ncalls"]))
        print(i, len(ints), round(time.time() - t0, 1), flush=True)
    a = np.array(rows, dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "serve_stream_oracle.npz"),
                        request=a[:, 0].astype(np.int64), position=a[:, 1].astype(np.int64),
                        stat=a[:, 2].astype(np.int64), term=a[:, 3].astype(np.int64),
                        value=a[:, 4], raw=a[:, 5], err=a[:, 6], neval=a[:, 7].astype(np.int64),
                        ncalls=a[:, 8].astype(np.int64), stream_seed=STREAM_SEED,
                        n_requests=N_REQUESTS, tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
