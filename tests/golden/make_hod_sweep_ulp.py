"""Writes tests/golden/hod_sweep8_oracle_ulp.npz: the oracle's 3-halo values
of the 8-HOD sweep fixture (tests/golden/make_hod_sweep_oracle.py: 8 prior
HOD sets x 55 triplets, tol 1e-4, maxEval 5e6) recomputed with its
sensitivity probe on (every exp, log, pow, sin, cos and erf result inside the
integrand moved by at most one unit in the last place, the n_bar(z)
integrals of each HOD set included: the seed is set before the HODs are),
seeds 1-8: how far the reference's algorithm itself moves each integral
under a one-ulp change of its math library.  Seeds given on the command line
are (re)computed and merged into an existing fixture.  About 10 minutes on 7
cores per seed.

Usage: OMP_NUM_THREADS=7 python tests/golden/make_hod_sweep_ulp.py [seed ...]
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

SEEDS = (1, 2, 3, 4, 5, 6, 7, 8)


def main():
    seeds = tuple(int(a) for a in sys.argv[1:]) or SEEDS
    path = os.path.join(HERE, "hod_sweep8_oracle_ulp.npz")
    have = {}
    if os.path.exists(path):
        old = np.load(path)
        have = {int(seed): old["value"][s] for s, seed in enumerate(old["seeds"])}
    O.build()
    t = T.make_tables()
    g = np.load(os.path.join(HERE, "hod_sweep8_oracle.npz"))
    hods = g["hods"]
    trips = T.triplets_55()
    t0 = time.time()
    try:
        for seed in seeds:
            O.set_ulp_seed(seed)
            out = np.zeros(len(hods) * len(trips))
            k = 0
            for h in range(len(hods)):
                m = O.Model(t)
                m.set_hods(hods[h], hods[h])
                for th in trips:
                    r = m.eval_term(0, 3, th * T.ARCMIN_IN_RAD, 1, 2, 1e-4, 5000000)
                    out[k] = r["value"]
                    k += 1
                print(seed, h, round(time.time() - t0, 1), flush=True)
            have[seed] = out
    finally:
        O.set_ulp_seed(0)
    keys = sorted(have)
    np.savez_compressed(path, seeds=np.array(keys), value=np.stack([have[k] for k in keys]),
                        tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
