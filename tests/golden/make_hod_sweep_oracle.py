"""Writes tests/golden/hod_sweep8_oracle.npz: the oracle's values for the
first 8 HOD vectors of the batched fit sweep (256 draws from the uniform
prior with seed 20220406, g3lhalo_amd.tables.hod_prior_draws) x the 55
triplets (a_i, a_j, a_j) x NNMap 1h/2h/3h, types (1,2), hod1 = hod2 as two
distinct objects, tol 1e-4, maxEval 5e6, on the 128-bin synthetic tables:
1320 integrals at the sweep's full shape per HOD set.  About 45 minutes on 8
cores.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_hod_sweep_oracle.py
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

N_HODS = 8
SWEEP_SEED = 20220406
SWEEP_DRAWS = 256


def main():
    O.build()
    t = T.make_tables()
    m = O.Model(t)
    hods = T.hod_prior_draws(SWEEP_DRAWS, seed=SWEEP_SEED)[:N_HODS]
    trips = T.triplets_55()
    n = N_HODS * len(trips) * 3
    out = np.zeros((n, 4))
    ncalls = np.zeros(n, dtype=np.int64)
    t0 = time.time()
    k = 0
    for h in range(N_HODS):
        m.set_hods(hods[h], hods[h])
        for th in trips:
            for term in (1, 2, 3):
                r = m.eval_term(0, term, th * T.ARCMIN_IN_RAD, 1, 2, 1e-4, 5000000)
                out[k] = (r["value"], r["raw"], r["err"], r["neval"])
                ncalls[k] = r["ncalls"]
                k += 1
        print(h, round(time.time() - t0, 1), flush=True)
    np.savez_compressed(os.path.join(HERE, "hod_sweep8_oracle.npz"), hods=hods,
                        value=out[:, 0], raw=out[:, 1], err=out[:, 2],
                        neval=out[:, 3].astype(np.int64), ncalls=ncalls,
                        tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
