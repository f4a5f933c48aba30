"""The gfx950 path against the committed vectors of the reference's own
adaptive cubature (tests/golden/cubature_ref.json): plain path with the
reference's evaluation count, batched path within 1e-4 relative per term,
device densities against the reference-driven n_bar(z)."""
import json
import os

import numpy as np
import pytest

from conftest import ARCMIN, ROOT, T, say

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "cubature_ref.json")))
TOL = 1e-4


def fh(s):
    return float.fromhex(s)


@pytest.fixture(scope="module")
def gctx(tables):
    from g3lhalo_amd import _lib
    assert T.checksum(tables) == GOLD["tables_checksum"]
    say("golden: context + two HOD sets")
    c = _lib.Context(tables)
    two = c.hodset(T.FIDUCIAL_HOD, GOLD["hod2"])
    same = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD, identical=True)
    yield c, two, same
    c.close()


def _req(g, sid):
    from g3lhalo_amd import _lib
    th = np.array(g["theta_arcmin"]) * ARCMIN
    return _lib.make_request(g["stat"], g["term"], th, hodset=sid, type1=g["types"][0],
                             type2=g["types"][1], reltol=g["reltol"], maxeval=g["maxeval"])


def test_densities_vs_reference_vectors(gctx):
    c, two, _ = gctx
    d = GOLD["densities"]
    for got, key in zip(c.densities(two), ("rho_bar", "n_bar1", "n_bar2")):
        ref = np.array([fh(x) for x in d[key]])
        assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12


@pytest.mark.parametrize("i", range(len(GOLD["terms"])))
def test_plain_vs_reference_vectors(gctx, i):
    c, two, same = gctx
    g = GOLD["terms"][i]
    say(f"golden plain case {i}: {g['stat']} {g['term']} {g['theta_arcmin']} {g['types']}")
    r = c.eval_plain(_req(g, same if g["stat"] == 1 else two))
    v = fh(g["value"])
    print(r.value, v, abs(r.value - v) / abs(v), r.neval, g["neval"])
    assert r.neval == g["neval"]
    assert abs(r.value - v) / abs(v) < TOL


def test_batch_vs_reference_vectors(gctx):
    c, two, same = gctx
    say("golden batch: all cases in one batch")
    reqs = [_req(g, same if g["stat"] == 1 else two) for g in GOLD["terms"]]
    for g, r in zip(GOLD["terms"], c.eval_batch(reqs)):
        v = fh(g["value"])
        d = abs(r.value - v) / abs(v)
        print(g["stat"], g["term"], r.value, v, d, r.neval, g["neval"])
        assert d < TOL
        pts = {1: 93, 2: 149, 3: 241}[g["term"]]
        assert abs(int(r.neval) - g["neval"]) <= 8 * pts
