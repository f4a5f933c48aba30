"""Writes tests/golden/nnmap550_oracle_ulp.npz: the oracle's 3-halo values of
the 1650-integral NNMap workload (550 aperture triplets, fiducial HOD, types
(1,2), tol 1e-4, maxEval 5e6) recomputed with its sensitivity probe on: every
exp, log, pow, sin, cos and erf result inside the integrand moved by at most
one unit in the last place (seeds 1 and 2), which is what a second, equally
accurate math library does to the same algorithm -- except the
nearest-index chains, which any implementation that matches the reference
has to reproduce bit for bit.

Every 3-halo integral of this workload stops at the evaluation cap
unconverged, so its value depends on which near-tied regions were split.  The
spread between these values and the unperturbed oracle's
(tests/golden/nnmap550_oracle.npz) is the scale on which any second
implementation of the reference's algorithm -- the reference's own build
against another libm included -- can agree with it.  About 35 minutes on 8
cores.

Usage: OMP_NUM_THREADS=8 python tests/golden/make_nnmap550_ulp.py
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

SEEDS = (1, 2)


def main():
    O.build()
    t = T.make_tables()
    m = O.Model(t)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    trips = T.triplets_550()
    out = np.zeros((len(SEEDS), len(trips), 3))
    t0 = time.time()
    try:
        for s, seed in enumerate(SEEDS):
            O.set_ulp_seed(seed)
            for i, th in enumerate(trips):
                r = m.eval_term(0, 3, th * T.ARCMIN_IN_RAD, 1, 2, 1e-4, 5000000)
                out[s, i] = (r["value"], r["err"], r["neval"])
                if i % 50 == 0:
                    print(seed, i, round(time.time() - t0, 1), flush=True)
    finally:
        O.set_ulp_seed(0)
    np.savez_compressed(os.path.join(HERE, "nnmap550_oracle_ulp.npz"), seeds=np.array(SEEDS),
                        value=out[:, :, 0], err=out[:, :, 1], neval=out[:, :, 2].astype(np.int64),
                        tables_checksum=T.checksum(t))


if __name__ == "__main__":
    main()
