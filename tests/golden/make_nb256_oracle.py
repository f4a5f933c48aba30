"""Writes tests/golden/nnmap550_nb256_oracle.npz: the oracle's results for a
sample of the 550-triplet NNMap workload on 256-bin tables (every table on a
256-point grid: the reference keeps one bin count for all of them,
src/NNMap_Model.cu:58-61): triplets 0, 183, 366 and 549 of the grid x 1h/2h/3h,
fiducial HOD, types (1,2), tol 1e-4, maxEval 5e6.  A few minutes on 8 cores.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_nb256_oracle.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from g3lhalo_amd import tables as T  # noqa: E402

NBINS = 256
SAMPLE = (0, 183, 366, 549)


def main():
    O.build()
    t = T.make_tables(nbins=NBINS)
    m = O.Model(t)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    trips = T.triplets_550()
    rows = []
    for i in SAMPLE:
        for term in (1, 2, 3):
            r = m.eval_term(0, term, trips[i] * T.ARCMIN_IN_RAD, 1, 2, 1e-4, 5000000)
            rows.append((i, term, r["value"], r["raw"], r["err"], r["neval"], r["ncalls"]))
            print(i, term, r["value"], flush=True)
    a = np.array(rows)
    np.savez_compressed(os.path.join(HERE, "nnmap550_nb256_oracle.npz"),
                        triplet=a[:, 0].astype(np.int64), term=a[:, 1].astype(np.int64),
                        value=a[:, 2], raw=a[:, 3], err=a[:, 4], neval=a[:, 5].astype(np.int64),
                        ncalls=a[:, 6].astype(np.int64), nbins=NBINS,
                        tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
