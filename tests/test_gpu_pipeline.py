"""The reference's end-to-end flow (scripts_examples/calculateNNMap.sh,
calculateNMapMap.sh): doPrecomputations.x writes the tables, then
calculateNNMap.x / calculateNMapMap.x read them and print the aperture
statistics.  Both steps run on the device here, and the drivers also run on
the tables the reference's own doPrecomputations.x wrote for the same grid
(tests/golden/precompute_pipeline16/).  Each printed term is compared with the
oracle's value of the same integral over exactly the tables the driver read
(the C++ ingest's resampling, dumped by tests/cxx/ingest_probe.cpp), and the
two table sources must give the same statistics.

The grid runs over redshifts 0.01-2.01 with 16 bins.  The precomputation
prints abscissae with six significant digits, and src/Function.cpp:55-57
returns the default value 0 for a redshift above the last printed abscissa;
on this grid the last one prints above its exact value, so every bin reads
back (on 0.001-2.0 the last bin of g, w and dw/dz would read as 0, in the
reference as here)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ARCMIN, ROOT, T, say

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "g3lhalo_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden", "precompute_pipeline16")
GRID = dict(zmin=0.01, zmax=2.01, kmin=1e-4, kmax=1e4, mmin=1e10, mmax=1e17, nbins=16)
OUTPUTS = ("g.dat", "w.dat", "dwdz.dat", "hmf.dat", "pk.dat", "bh.dat", "conc.dat")
TOL = 1e-4 + 5e-6                  # batched term tolerance + six-digit printing
NNMAP_TOL = (1e-3, 1e-3, 1e-4)     # src/NNMap_Model.cu:318,343,368
NMAPMAP_TOL = (1e-3, 1e-3, 1e-3)   # src/NMapMap_Model.cu:186,211,234
THETAS = (4.0, 4.0, 4.0)
# the tables of the two precomputations agree to their six printed digits;
# the statistics then agree to the terms' own tolerances
CROSS_TOL = 2e-3


def _grid_args():
    g = GRID
    return [repr(g["zmin"]), repr(g["zmax"]), repr(g["kmin"]), repr(g["kmax"]),
            repr(g["mmin"]), repr(g["mmax"]), str(g["nbins"])]


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("probe") / "ingest_probe")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "ingest_probe.cpp"), "-o", out,
                        "-L", os.path.join(ROOT, "g3lhalo_amd"), "-l:libg3lhalo.so",
                        "-l:libg3l_hip.so", "-Wl,-rpath," + os.path.join(ROOT, "g3lhalo_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _lens_nz(path):
    z = np.linspace(0.0, 2.1, 211)
    n = np.exp(-0.5 * ((z - 0.5) / 0.15) ** 2)
    n /= n.sum() * (z[1] - z[0])
    with open(path, "w") as f:
        for a, b in zip(z, n):
            f.write(f"{a:.17g} {b:.17g}\n")


def _hod(path, h):
    f_, alpha, mth, sig, mp, beta = (float(v) for v in h)
    with open(path, "w") as f:
        f.write(f"M_min {mth!r}\nsigma {sig!r}\nM' {mp!r}\nbeta {beta!r}\nalpha {alpha!r}\n"
                f"f {f_!r}\n")


@pytest.fixture(scope="module", params=("device", "reference"))
def workdir(request, tmp_path_factory):
    d = tmp_path_factory.mktemp(f"pipeline_{request.param}")
    for f in ("cosmo.param", "nz_sources.dat"):
        shutil.copy(os.path.join(GOLD, f), d / f)
    if request.param == "device":
        say("doPrecomputations.x pipeline16")
        subprocess.run([os.path.join(BIN, "doPrecomputations.x"), "cosmo.param",
                        "nz_sources.dat", repr(GRID["zmin"]), repr(GRID["zmax"]),
                        repr(GRID["mmin"]), repr(GRID["mmax"]), repr(GRID["kmin"]),
                        repr(GRID["kmax"]), str(GRID["nbins"]), *OUTPUTS],
                       cwd=d, check=True, timeout=120)
    else:
        for f in OUTPUTS:
            shutil.copy(os.path.join(GOLD, f), d / f)
    _lens_nz(d / "pz_lenses.dat")
    _hod(d / "hod1.param", T.FIDUCIAL_HOD)
    _hod(d / "hod2.param", T.hod_prior_draws(1, seed=7)[0])
    with open(d / "thetas.dat", "w") as f:
        f.write(" ".join(repr(v) for v in THETAS) + "\n")
    return request.param, d


def _tables(probe, d):
    r = subprocess.run([probe, "dump", *(_grid_args()[i] for i in (6, 0, 1, 2, 3, 4, 5)),
                        *(str(d / f) for f in ("g.dat", "pz_lenses.dat", "pz_lenses.dat",
                                               "w.dat", "dwdz.dat", "hmf.dat", "pk.dat",
                                               "bh.dat", "conc.dat"))],
                       capture_output=True, text=True, check=True)
    a = {ln.split()[0]: np.array(ln.split()[1:], float) for ln in r.stdout.splitlines()}
    n = GRID["nbins"]
    for k in ("g", "p_lens1", "w", "dwdz"):
        assert a[k].shape == (n,) and np.all(np.isfinite(a[k]))
        assert a[k][-1] != 0.0, k
    for k in ("hmf", "P_lin", "b_h", "concentration"):
        assert a[k].shape == (n * n,) and np.all(np.isfinite(a[k]))
    c = T.Cosmo(Omega_m=0.3, Omega_Lambda=0.7, Omega_b=0.045, h=0.7, sigma8=0.8, ns=0.96)
    return T.Tables(nbins=n, zmin=GRID["zmin"], zmax=GRID["zmax"], kmin=GRID["kmin"],
                    kmax=GRID["kmax"], mmin=GRID["mmin"], mmax=GRID["mmax"], cosmo=c,
                    g=a["g"], p_lens1=a["p_lens1"], p_lens2=a["p_lens2"], w=a["w"],
                    dwdz=a["dwdz"], hmf=a["hmf"], P_lin=a["P_lin"], b_h=a["b_h"],
                    concentration=a["concentration"])


def _run(exe, args, cwd):
    say(f"driver {exe} in {cwd}")
    r = subprocess.run([os.path.join(BIN, exe), *args], capture_output=True, text=True,
                       timeout=110, cwd=cwd)
    assert r.returncode == 0, r.stderr
    rows = np.array([list(map(float, ln.split())) for ln in r.stdout.strip().splitlines()])
    assert rows.shape == (1, 6), r.stdout
    assert np.allclose(rows[0, :3], THETAS)
    return rows[0]


def _check(row, model, stat, tols):
    th = row[:3] * ARCMIN
    for term in (1, 2, 3):
        o = model.eval_term(stat, term, th, type1=1, type2=2, reltol=tols[term - 1])["value"]
        got = row[2 + term]
        d = abs(got - o) / abs(o)
        print(f"stat={stat} term={term} {got:.6g} oracle {o:.6g} rel {d:.2e}")
        assert np.isfinite(got) and d < TOL


_rows = {}


def test_precompute_then_nnmap(probe, workdir):
    import oracle as O
    src, d = workdir
    files = ["g.dat", "pz_lenses.dat", "pz_lenses.dat", "w.dat", "dwdz.dat", "hmf.dat",
             "pk.dat", "bh.dat", "conc.dat", "hod1.param", "hod2.param", "0", "0",
             "thetas.dat"]
    row = _run("calculateNNMap.x", ["cosmo.param", *_grid_args(), *files], d)
    m = O.Model(_tables(probe, d))
    m.set_hods(T.FIDUCIAL_HOD, T.hod_prior_draws(1, seed=7)[0], A=0.0, eps=0.0,
               identical=False)
    _check(row, m, 0, NNMAP_TOL)
    _rows[("nnmap", src)] = row


def test_precompute_then_nmapmap(probe, workdir):
    import oracle as O
    src, d = workdir
    files = ["g.dat", "pz_lenses.dat", "w.dat", "dwdz.dat", "hmf.dat", "pk.dat", "bh.dat",
             "conc.dat", "hod1.param", "thetas.dat"]
    row = _run("calculateNMapMap.x", ["cosmo.param", *_grid_args(), *files], d)
    m = O.Model(_tables(probe, d))
    m.set_hods(T.FIDUCIAL_HOD)
    _check(row, m, 1, NMAPMAP_TOL)
    _rows[("nmapmap", src)] = row


@pytest.mark.parametrize("stat", ("nnmap", "nmapmap"))
def test_device_and_reference_tables_agree(stat):
    if (stat, "device") not in _rows or (stat, "reference") not in _rows:
        pytest.skip("needs both pipeline runs of this module")
    a, b = _rows[(stat, "device")], _rows[(stat, "reference")]
    rel = np.abs(a[3:] - b[3:]) / np.abs(b[3:])
    print(f"{stat}: device-table vs reference-table terms rel {rel}")
    assert np.all(rel < CROSS_TOL)
