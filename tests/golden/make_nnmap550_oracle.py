"""Writes tests/golden/nnmap550_oracle.npz: the oracle's value, raw integral,
raw error estimate, evaluation count, number of integrand calls (refinement
rounds, the root evaluation included) and batch-size trace (points per call)
for every integral of the 1650-integral NNMap workload (550 aperture triplets
x 1h/2h/3h, fiducial HOD, types (1,2), tol 1e-4, maxEval 5e6) on the 128-bin
synthetic tables, for the full-size parity test of the device path.  About
an hour on 8 cores.

At the evaluation cap the count is fixed by arithmetic whatever regions were
split; the call count and the trace are what record the refinement sequence.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_nnmap550_oracle.py [stat]
(stat 0 = NNMap, the default; 1 = NMapMap writes nmapmap550_oracle.npz)
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

TRACE_CAP = 64


def main():
    stat = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    O.build()
    t = T.make_tables()
    m = O.Model(t)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    trips = T.triplets_550()
    n = len(trips) * 3
    out = np.zeros((n, 4))
    ncalls = np.zeros(n, dtype=np.int64)
    trace = np.zeros((n, TRACE_CAP), dtype=np.uint32)
    t0 = time.time()
    for i, th in enumerate(trips):
        for j, term in enumerate((1, 2, 3)):
            r = m.eval_term(stat, term, th * T.ARCMIN_IN_RAD, 1, 2, 1e-4, 5000000,
                            trace_cap=TRACE_CAP)
            k = 3 * i + j
            out[k] = (r["value"], r["raw"], r["err"], r["neval"])
            ncalls[k] = r["ncalls"]
            assert r["ncalls"] == len(r["trace"]) <= TRACE_CAP
            trace[k, :len(r["trace"])] = r["trace"]
        if i % 50 == 0:
            print(i, round(time.time() - t0, 1), flush=True)
    name = "nnmap550_oracle.npz" if stat == 0 else "nmapmap550_oracle.npz"
    np.savez_compressed(os.path.join(HERE, name), value=out[:, 0], raw=out[:, 1], err=out[:, 2],
                        neval=out[:, 3].astype(np.int64), ncalls=ncalls, trace=trace,
                        tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
