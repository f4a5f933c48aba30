"""Host-side checks of the reference-compatible C++ API's file ingest
(Function, Function2D, Params, Cosmology): files written in the reference's
ASCII formats read back to the exact table values."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, T


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    lib = os.path.join(ROOT, "g3lhalo_amd", "libg3lhalo.so")
    if not os.path.exists(lib):
        pytest.skip("libg3lhalo.so not built")
    out = str(tmp_path_factory.mktemp("probe") / "ingest_probe")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cxx", "ingest_probe.cpp"), "-o", out,
                        "-L", os.path.join(ROOT, "g3lhalo_amd"), "-l:libg3lhalo.so",
                        "-l:libg3l_hip.so", "-Wl,-rpath," + os.path.join(ROOT, "g3lhalo_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_reference_files_round_trip(probe, tmp_path):
    t = T.make_tables(nbins=16)
    T.write_reference_files(t, str(tmp_path))
    r = subprocess.run([probe, str(tmp_path), "16", repr(t.zmin), repr(t.zmax), repr(t.mmin),
                        repr(t.mmax)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = {ln.split()[0]: ln.split()[1:] for ln in r.stdout.strip().splitlines()}
    assert np.array_equal(np.array(lines["w"], float), t.w)
    assert np.array_equal(np.array(lines["hmf"], float), t.hmf)
    assert np.array_equal(np.array(lines["params"], float), np.array(T.FIDUCIAL_HOD)[[0, 1, 2, 3, 4, 5]])
    c = np.array(lines["cosmo"], float)
    assert c[0] == t.cosmo.Omega_m and c[2] == t.cosmo.H0
    assert c[3] == pytest.approx(3 * 70.0 ** 2 * 3.086e19 / 8 / 3.14159 / 1.327e11, rel=1e-15)
    assert float(lines["edge"][0]) == 0.0


def test_params_rejects_unknown_names(probe, tmp_path):
    t = T.make_tables(nbins=16)
    T.write_reference_files(t, str(tmp_path))
    # the reference's shipped hod files use suffixed names, which Params rejects
    with open(tmp_path / "hod1.param", "w") as f:
        f.write("M_min1 5e+13\nsigma_1 0.7\n")
    r = subprocess.run([probe, str(tmp_path), "16", repr(t.zmin), repr(t.zmax), repr(t.mmin),
                        repr(t.mmax)], capture_output=True, text=True)
    assert r.returncode == 1
    assert "not in the right format" in r.stderr


def test_driver_rejects_wrong_argument_count():
    exe = os.path.join(ROOT, "g3lhalo_amd", "bin", "calculateNNMap.x")
    if not os.path.exists(exe):
        pytest.skip("drivers not built")
    # checkCmdLine exits before any device use (src/helpers.cpp)
    r = subprocess.run([exe, "x"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1
    assert "Expected:23" in r.stderr and "Got:2" in r.stderr
