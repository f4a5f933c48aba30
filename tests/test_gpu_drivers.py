"""The drop-in drivers (bin/calculateNNMap.x, calculateNNMap_twoPop.x,
calculateNMapMap.x) run end to end on the device: reference-format input
files in, "theta1 theta2 theta3 1h 2h 3h" rows out, each term compared with
the oracle's value of the same integral.

The drivers print six significant digits (src/calculateNNMap.cu:148-154), so
the bar is the term tolerance of the batched path (1e-4 relative) plus the
printing's half unit in the sixth digit."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ARCMIN, ROOT, T, say

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "g3lhalo_amd", "bin")
TOL = 1e-4 + 5e-6
NNMAP_TOL = (1e-3, 1e-3, 1e-4)   # src/NNMap_Model.cu:318,343,368
NMAPMAP_TOL = (1e-3, 1e-3, 1e-3)  # src/NMapMap_Model.cu:186,211,234
A_CORR, EPS_CORR = 0.5, 0.1


@pytest.fixture(scope="module")
def files(tables, tmp_path_factory):
    d = str(tmp_path_factory.mktemp("drivers"))
    hod2 = tuple(T.hod_prior_draws(1, seed=7)[0])
    f = T.write_reference_files(tables, d, hods=(T.FIDUCIAL_HOD, hod2))
    for name, rows in (("thetas2.dat", [(4.0, 4.0, 4.0), (2.0, 6.0, 9.0)]),
                       ("thetas1.dat", [(5.0, 3.0, 8.0)])):
        with open(os.path.join(d, name), "w") as fh:
            for r in rows:
                fh.write(" ".join(repr(v) for v in r) + "\n")
        f[name] = os.path.join(d, name)
    return f, hod2


def _common(t, f):
    return [f["cosmo.param"], repr(t.zmin), repr(t.zmax), repr(t.kmin), repr(t.kmax),
            repr(t.mmin), repr(t.mmax), str(t.nbins), f["g.dat"]]


def _tables(f):
    return [f["w.dat"], f["dwdz.dat"], f["hmf.dat"], f["Plin.dat"], f["b_h.dat"],
            f["concentration.dat"]]


def _run(exe, args):
    say(f"driver {exe}")
    r = subprocess.run([os.path.join(BIN, exe), *args], capture_output=True, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stderr
    rows = [list(map(float, ln.split())) for ln in r.stdout.strip().splitlines()]
    assert all(len(x) == 6 for x in rows), r.stdout
    return np.array(rows)


def _nnmap_args(t, f, thetas):
    return [*_common(t, f), f["pz_lens1.dat"], f["pz_lens2.dat"], *_tables(f), f["hod1.param"],
            f["hod2.param"], repr(A_CORR), repr(EPS_CORR), f[thetas]]


def _check(rows, oracle, stat, tols, type1=1, type2=2):
    for row in rows:
        th = row[:3] * ARCMIN
        for term in (1, 2, 3):
            o = oracle.eval_term(stat, term, th, type1=type1, type2=type2,
                                 reltol=tols[term - 1])["value"]
            got = row[2 + term]
            d = abs(got - o) / abs(o)
            print(f"stat={stat} types=({type1},{type2}) term={term} theta={row[:3]} "
                  f"{got:.6g} oracle {o:.6g} rel {d:.2e}")
            assert d < TOL


@pytest.fixture(scope="module")
def two_pop_oracle(tables, files):
    import oracle as O
    m = O.Model(tables)
    m.set_hods(T.FIDUCIAL_HOD, files[1], A=A_CORR, eps=EPS_CORR, identical=False)
    return m


def test_calculate_nnmap(tables, files, two_pop_oracle):
    f, _ = files
    rows = _run("calculateNNMap.x", _nnmap_args(tables, f, "thetas2.dat"))
    assert rows.shape == (2, 6)
    assert np.allclose(rows[:, :3], [[4, 4, 4], [2, 6, 9]])
    _check(rows, two_pop_oracle, 0, NNMAP_TOL)


def test_calculate_nnmap_two_pop(tables, files, two_pop_oracle):
    f, _ = files
    rows = _run("calculateNNMap_twoPop.x", _nnmap_args(tables, f, "thetas1.dat"))
    assert rows.shape == (3, 6)
    for row, (a, b) in zip(rows, ((1, 2), (1, 1), (2, 2))):
        _check(row[None, :], two_pop_oracle, 0, NNMAP_TOL, a, b)


def test_calculate_nmapmap(tables, files):
    import oracle as O
    f, _ = files
    m = O.Model(tables)
    m.set_hods(T.FIDUCIAL_HOD)
    args = [*_common(tables, f), f["pz_lens1.dat"], *_tables(f), f["hod1.param"],
            f["thetas2.dat"]]
    rows = _run("calculateNMapMap.x", args)
    assert rows.shape == (2, 6)
    _check(rows, m, 1, NMAPMAP_TOL)

