"""The correctly rounded 10^x and cos(x) of g3lhalo_amd/csrc/crmath.h (the
NNMap 3-halo integrand's l1, l2 and cos(phi)) against mpmath at 250 bits:
every result must be the double nearest the exact value, on the arguments
the integrands use (log10 l in [-1, 6], log10 m in [10, 17.5], phi in
[0, 2 x 3.14159]) and on exact special points.  The host build compiles the
same source as the device (IEEE +, -, x, FMA, rint and ldexp only), with
contraction off as in the device build."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

mpmath = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cr") / "crmath_probe")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                        os.path.join(ROOT, "tests", "cxx", "crmath_probe.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, op, xs):
    inp = "".join(f"{op} {float(x).hex()}\n" for x in xs)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
    return [float.fromhex(v) for v in out.stdout.split()]


def test_exp10_correctly_rounded(probe):
    mpmath.mp.prec = 250
    rng = np.random.default_rng(101)
    xs = np.concatenate([rng.uniform(-1, 6, 6000), rng.uniform(10, 17.5, 4000),
                         np.arange(-10, 18, 0.5)])
    got = _run(probe, "e", xs)
    bad = [x for x, v in zip(xs, got) if v != float(mpmath.power(10, mpmath.mpf(float(x))))]
    assert not bad, bad[:5]
    assert _run(probe, "e", [0.0, 1.0, 2.0, 3.0]) == [1.0, 10.0, 100.0, 1000.0]


def test_cos_correctly_rounded(probe):
    mpmath.mp.prec = 250
    rng = np.random.default_rng(102)
    xs = np.concatenate([rng.uniform(0, 2 * 3.14159, 10000),
                         [0.0, 1e-12, math.pi / 4, math.pi / 2, 3.14159, 2 * 3.14159]])
    got = _run(probe, "c", xs)
    bad = [x for x, v in zip(xs, got) if v != float(mpmath.cos(mpmath.mpf(float(x))))]
    assert not bad, bad[:5]
