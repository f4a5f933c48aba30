"""Writes tests/golden/hod_sweep8_oracle_ulp.npz: the oracle's 3-halo values
of the 8-HOD sweep fixture (tests/golden/make_hod_sweep_oracle.py: 8 prior
HOD sets x 55 triplets, tol 1e-4, maxEval 5e6) recomputed with its
sensitivity probe on (every exp, log, pow, sin, cos and erf result inside the
integrand moved by at most one unit in the last place), seeds 1-4: the
integrals whose value the reference's algorithm itself does not fix to 1e-4
under a one-ulp change of its math library.  About 35 minutes on 8 cores.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_hod_sweep_ulp.py
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

SEEDS = (1, 2, 3, 4)


def main():
    O.build()
    t = T.make_tables()
    g = np.load(os.path.join(HERE, "hod_sweep8_oracle.npz"))
    hods = g["hods"]
    trips = T.triplets_55()
    out = np.zeros((len(SEEDS), len(hods) * len(trips)))
    t0 = time.time()
    try:
        for s, seed in enumerate(SEEDS):
            O.set_ulp_seed(seed)
            k = 0
            for h in range(len(hods)):
                m = O.Model(t)
                m.set_hods(hods[h], hods[h])
                for th in trips:
                    r = m.eval_term(0, 3, th * T.ARCMIN_IN_RAD, 1, 2, 1e-4, 5000000)
                    out[s, k] = r["value"]
                    k += 1
                print(seed, h, round(time.time() - t0, 1), flush=True)
    finally:
        O.set_ulp_seed(0)
    np.savez_compressed(os.path.join(HERE, "hod_sweep8_oracle_ulp.npz"), seeds=np.array(SEEDS),
                        value=out, tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
