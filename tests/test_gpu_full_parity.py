"""Full-size parity: the 1650-integral NNMap and NMapMap workloads (550
aperture triplets x 1h/2h/3h, tol 1e-4, maxEval 5e6) on the device against the
oracle's values for every integral (tests/golden/nnmap550_oracle.npz,
tests/golden/nmapmap550_oracle.npz).

Every integral of this workload stops at the evaluation cap unconverged; its
value depends on which regions were split, and a last-place difference in an
integrand value can change that choice.  Every NMapMap integral and the
NNMap 1-halo and 2-halo integrals refine identically (equal evaluation counts,
values equal to rounding); the 7-dimensional NNMap 3-halo integrals, whose
error estimates are 3 % to several times their values, mostly do.  Where they
do not, the difference is of the size the oracle itself shows when its math
library moves by one unit in the last place (tests/golden/
nnmap550_oracle_ulp.npz: the same few integrals move, by up to ~5e-4), and it
stays far inside the integral's own error estimate."""
import os

import numpy as np
import pytest

from conftest import ROOT, T, say

pytestmark = pytest.mark.gpu

GOLD = {0: os.path.join(ROOT, "tests", "golden", "nnmap550_oracle.npz"),
        1: os.path.join(ROOT, "tests", "golden", "nmapmap550_oracle.npz")}
ULP = os.path.join(ROOT, "tests", "golden", "nnmap550_oracle_ulp.npz")


@pytest.mark.parametrize("stat", [0, 1])
def test_550_against_oracle(tables, stat):
    from g3lhalo_amd import _lib
    g = np.load(GOLD[stat])
    assert str(g["tables_checksum"]) == T.checksum(tables)
    c = _lib.Context(tables)
    try:
        sid = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        reqs = [_lib.make_request(stat, term, th * T.ARCMIN_IN_RAD, hodset=sid, reltol=1e-4,
                                  maxeval=5000000)
                for th in T.triplets_550() for term in (1, 2, 3)]
        say(f"full 550-triplet batch, stat {stat}")
        out = c.eval_batch(reqs)
    finally:
        c.close()
    v = np.array([o.value for o in out])
    ne = np.array([o.neval for o in out])
    d = np.abs(v - g["value"]) / np.abs(g["value"])
    err_value = np.abs(g["err"] * g["value"] / g["raw"])   # the oracle's error in value units
    for term in (1, 2, 3):
        dt = d[term - 1::3]
        print(f"term {term}: median {np.median(dt):.2e} p99 {np.quantile(dt, 0.99):.2e} "
              f"max {dt.max():.2e} above 1e-4: {(dt > 1e-4).sum()}")
    assert np.array_equal(ne, g["neval"])
    assert d[0::3].max() < 1e-10 and d[1::3].max() < 1e-10
    if stat == 1:
        # NMapMap's 3-halo integrals refine identically too
        assert d[2::3].max() < 1e-10
        return
    d3 = d[2::3]
    # the scale set by the reference's algorithm itself: the oracle's 3-halo
    # values under one-ulp perturbations of its math library (two seeds)
    u = np.load(ULP)
    assert str(u["tables_checksum"]) == T.checksum(tables)
    du = np.abs(u["value"] - g["value"][2::3]) / np.abs(g["value"][2::3])
    n_ulp = int((du > 1e-4).sum(axis=1).max())
    print(f"NNMap 3h: device above 1e-4: {(d3 > 1e-4).sum()}, max {d3.max():.2e}; oracle under "
          f"1-ulp perturbation: above 1e-4 at most {n_ulp}, max {du.max():.2e}")
    assert (d3 > 1e-4).sum() <= 3 * max(n_ulp, 1)
    assert d3.max() <= 2 * du.max()
    assert np.median(d3) < 1e-7
    assert np.all(np.abs(v - g["value"]) < 0.1 * err_value)
