"""The 256-bin variant of the 550-triplet NNMap workload: every table on a
256-point grid (the reference's one bin count for all tables,
src/NNMap_Model.cu:58-61), a sample of integrals against the oracle's values
on the same tables (tests/golden/nnmap550_nb256_oracle.npz): the 1- and
2-halo terms to rounding with the oracle's integrand-call counts, the 3-halo
term inside its error estimate."""
import os

import numpy as np
import pytest

from conftest import ROOT, T, say

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "nnmap550_nb256_oracle.npz")


def test_nb256_sample_matches_oracle():
    from g3lhalo_amd import _lib
    g = np.load(GOLD)
    t = T.make_tables(nbins=int(g["nbins"]))
    assert str(g["tables_checksum"]) == T.checksum(t)
    trips = T.triplets_550()
    c = _lib.Context(t)
    try:
        sid = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        reqs = [_lib.make_request(0, int(term), trips[int(i)] * T.ARCMIN_IN_RAD, hodset=sid,
                                  reltol=1e-4, maxeval=5000000)
                for i, term in zip(g["triplet"], g["term"])]
        say(f"nb256: {len(reqs)} integrals")
        out = c.eval_batch(reqs)
    finally:
        c.close()
    for k, r in enumerate(out):
        d = abs(r.value - g["value"][k]) / abs(g["value"][k])
        print(f"triplet {g['triplet'][k]} term {g['term'][k]}: rel {d:.2e}, rounds {r.rounds} / "
              f"{g['ncalls'][k]}")
        assert r.neval == g["neval"][k]
        if g["term"][k] == 3:
            assert abs(r.raw - g["raw"][k]) < 0.1 * g["err"][k]
        else:
            assert d < 1e-10 and r.rounds == g["ncalls"][k]
