"""The oracle's own sensitivity on the full-size NNMap 3-halo workload
(tests/golden/nnmap550_oracle_ulp.npz, written by
tests/golden/make_nnmap550_ulp.py): the 3-halo integrals recomputed with the
elementary functions moved by up to one unit in the last place.  Evaluation
counts stay the same, most values move by rounding only, and a few of the
capped, unconverged integrals move by more than 1e-4 -- the scale against
which the device's values are judged (tests/test_gpu_full_parity.py).  One
entry of each fixture is recomputed here, so the fixtures stay tied to the
current oracle."""
import os

import numpy as np

from conftest import ROOT, T

import oracle as O

GOLD = os.path.join(ROOT, "tests", "golden", "nnmap550_oracle.npz")
ULP = os.path.join(ROOT, "tests", "golden", "nnmap550_oracle_ulp.npz")


def test_ulp_spread_of_the_3h_workload(tables):
    g, u = np.load(GOLD), np.load(ULP)
    assert str(g["tables_checksum"]) == T.checksum(tables) == str(u["tables_checksum"])
    v0 = g["value"][2::3]
    for s in range(len(u["seeds"])):
        assert np.array_equal(u["neval"][s], g["neval"][2::3])
    du = np.abs(u["value"] - v0) / np.abs(v0)
    moved = (du > 1e-4).sum(axis=1)
    assert len(u["seeds"]) >= 4
    assert 1 <= moved.max() <= 20
    assert du.max() < 1e-2
    assert np.median(du) < 1e-7


def test_ulp_fixture_nbar_seed_recomputes(tables):
    """A seed that perturbs the n_bar(z) integrals as well: the seed is set
    before the HOD set is."""
    g, u = np.load(GOLD), np.load(ULP)
    s = int(np.flatnonzero(u["nbar_perturbed"])[0])
    du = np.abs(u["value"][s] - g["value"][2::3]) / np.abs(g["value"][2::3])
    i = int(np.argmax(du))
    th = T.triplets_550()[i] * T.ARCMIN_IN_RAD
    try:
        O.set_ulp_seed(int(u["seeds"][s]))
        m = O.Model(tables)
        m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        r = m.eval_term(0, 3, th, 1, 2, 1e-4, 5000000)
    finally:
        O.set_ulp_seed(0)
    assert r["value"] == u["value"][s][i] and r["neval"] == u["neval"][s][i]


def test_ulp_fixture_entries_recompute(tables):
    g, u = np.load(GOLD), np.load(ULP)
    trips = T.triplets_550()
    du = np.abs(u["value"][0] - g["value"][2::3]) / np.abs(g["value"][2::3])
    i = int(np.argmax(du))          # the entry that moves most under seed 1
    m = O.Model(tables)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    th = trips[i] * T.ARCMIN_IN_RAD
    r0 = m.eval_term(0, 3, th, 1, 2, 1e-4, 5000000)
    assert r0["value"] == g["value"][3 * i + 2] and r0["neval"] == g["neval"][3 * i + 2]
    try:
        O.set_ulp_seed(int(u["seeds"][0]))
        r1 = m.eval_term(0, 3, th, 1, 2, 1e-4, 5000000)
    finally:
        O.set_ulp_seed(0)
    assert r1["value"] == u["value"][0][i] and r1["neval"] == u["neval"][0][i]


SWEEP = os.path.join(ROOT, "tests", "golden", "hod_sweep8_oracle.npz")
SWEEP_ULP = os.path.join(ROOT, "tests", "golden", "hod_sweep8_oracle_ulp.npz")


def test_ulp_spread_of_the_sweep_3h(tables):
    """The same probe over the 8-HOD sweep fixture's 440 3-halo integrals
    (tests/golden/hod_sweep8_oracle_ulp.npz): prior HOD sets move more of
    them than the fiducial set does."""
    g, u = np.load(SWEEP), np.load(SWEEP_ULP)
    assert str(g["tables_checksum"]) == T.checksum(tables) == str(u["tables_checksum"])
    v0 = g["value"][2::3]
    assert u["value"].shape == (len(u["seeds"]), len(v0)) and len(u["seeds"]) >= 4
    du = np.abs(u["value"] - v0) / np.abs(v0)
    moved = (du > 1e-4).sum(axis=1)
    assert 1 <= moved.min() and moved.max() <= 60
    assert du.max() < 1e-2
    assert np.median(du) < 1e-5


def test_sweep_ulp_fixture_entry_recomputes(tables):
    g, u = np.load(SWEEP), np.load(SWEEP_ULP)
    trips = T.triplets_55()
    du = np.abs(u["value"][0] - g["value"][2::3]) / np.abs(g["value"][2::3])
    i = int(np.argmax(du))
    h, t = divmod(i, len(trips))
    th = trips[t] * T.ARCMIN_IN_RAD
    try:
        # the fixture's seeds perturb each HOD set's n_bar(z) integrals too
        O.set_ulp_seed(int(u["seeds"][0]))
        m = O.Model(tables)
        m.set_hods(g["hods"][h], g["hods"][h])
        r1 = m.eval_term(0, 3, th, 1, 2, 1e-4, 5000000)
    finally:
        O.set_ulp_seed(0)
    assert r1["value"] == u["value"][0][i]
