"""The table-precomputation fixtures (tests/golden/precompute_*/, written by
tests/golden/make_precompute_golden.py from the reference's own
doPrecomputations.x): their layout, and -- where oracle/_ref holds the
reference build -- that a fresh run of the reference reproduces them byte for
byte."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(ROOT, "oracle", "_ref", "doPrecomputations.x")
OUTPUTS = ("g.dat", "w.dat", "dwdz.dat", "hmf.dat", "pk.dat", "bh.dat", "conc.dat")
CONFIGS = ("fiducial16", "example8", "pipeline16")


def load(name):
    d = os.path.join(GOLD, f"precompute_{name}")
    cfg = json.load(open(os.path.join(d, "config.json")))
    return d, cfg


@pytest.mark.parametrize("name", CONFIGS)
def test_fixture_layout(name):
    d, cfg = load(name)
    n = cfg["nbins"]
    z = cfg["zmin"] + np.arange(n) * (cfg["zmax"] - cfg["zmin"]) / n
    for f in ("g.dat", "w.dat", "dwdz.dat"):
        a = np.loadtxt(os.path.join(d, f))
        assert a.shape == (n, 2)
        assert np.allclose(a[:, 0], z, rtol=1e-5)
    for f, lo, hi in (("hmf.dat", "mmin", "mmax"), ("bh.dat", "mmin", "mmax"),
                      ("conc.dat", "mmin", "mmax"), ("pk.dat", "kmin", "kmax")):
        a = np.loadtxt(os.path.join(d, f))
        assert a.shape == (n * n, 3)
        x = np.exp(np.log(cfg[lo]) + np.arange(n) * np.log(cfg[hi] / cfg[lo]) / n)
        assert np.allclose(a[:n, 1], x, rtol=1e-5)
        assert np.allclose(a[::n, 0], z, rtol=1e-5)
    w = np.loadtxt(os.path.join(d, "w.dat"))[:, 1]
    assert np.all(np.diff(w) > 0)          # comoving distance grows with z
    hmf = np.loadtxt(os.path.join(d, "hmf.dat"))[:, 2].reshape(n, n)
    assert np.all(hmf > 0) and np.all(np.diff(hmf, axis=1) < 0)   # n(m) falls with mass


@pytest.mark.parametrize("name", CONFIGS)
def test_reference_build_reproduces_fixture(name, tmp_path):
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/doPrecomputations.x not built (reference sources absent)")
    d, cfg = load(name)
    for f in ("cosmo.param", "nz_sources.dat"):
        shutil.copy(os.path.join(d, f), tmp_path / f)
    args = [REF, "cosmo.param", "nz_sources.dat", repr(cfg["zmin"]), repr(cfg["zmax"]),
            repr(cfg["mmin"]), repr(cfg["mmax"]), repr(cfg["kmin"]), repr(cfg["kmax"]),
            str(cfg["nbins"]), *OUTPUTS]
    subprocess.run(args, cwd=tmp_path, check=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
    for f in OUTPUTS:
        assert open(tmp_path / f).read() == open(os.path.join(d, f)).read(), f
