"""The oracle against the committed vectors of the reference's own adaptive
cubature (tests/golden/cubature_ref.json, written by
tests/golden/make_golden.py from src/hcubature.cpp): bit-identical values,
errors, evaluation counts and batch traces, without the reference sources."""
import json
import os

import numpy as np
import pytest

from conftest import ARCMIN, ROOT, T

import oracle as O

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "cubature_ref.json")))


def fh(s):
    return float.fromhex(s)


@pytest.fixture(scope="module")
def models(tables):
    assert T.checksum(tables) == GOLD["tables_checksum"]
    two = O.Model(tables)
    two.set_hods(T.FIDUCIAL_HOD, GOLD["hod2"])
    same = O.Model(tables)
    same.set_hods(T.FIDUCIAL_HOD, identical=True)
    return two, same


@pytest.mark.parametrize("i", range(len(GOLD["genz"])))
def test_genz_vs_reference_vectors(i):
    g = GOLD["genz"][i]
    val, err, ne, nc = O.genz(g["family"], [fh(x) for x in g["a"]], [fh(x) for x in g["u"]],
                              [fh(x) for x in g["lo"]], [fh(x) for x in g["hi"]],
                              maxeval=g["maxeval"], reltol=g["reltol"])
    assert val == fh(g["val"]) and err == fh(g["err"])
    assert ne == g["neval"] and nc == g["ncalls"]


@pytest.mark.parametrize("i", range(len(GOLD["terms"])))
def test_terms_vs_reference_vectors(models, i):
    g = GOLD["terms"][i]
    two, same = models
    m = same if g["stat"] == 1 else two
    r = m.eval_term(g["stat"], g["term"], np.array(g["theta_arcmin"]) * ARCMIN, *g["types"],
                    reltol=g["reltol"], maxeval=g["maxeval"], trace_cap=4096)
    assert r["raw"] == fh(g["raw"]) and r["err"] == fh(g["err"])
    assert r["value"] == fh(g["value"])
    assert r["neval"] == g["neval"] and r["ncalls"] == g["ncalls"]
    assert list(r["trace"]) == g["trace"]


def test_densities_vs_reference_vectors(models):
    two, _ = models
    d = GOLD["densities"]
    for got, key in zip(two.densities(), ("rho_bar", "n_bar1", "n_bar2")):
        assert np.array_equal(got, np.array([fh(x) for x in d[key]]))
