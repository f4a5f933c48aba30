"""The FP64 policy's u_NFW arithmetic, built for the host from the same
source and checked against mpmath:

* sincos_mm (fdlibm's minimax kernels after the Cody-Waite reduction) within
  two units in the last place of sin and cos at 40 digits over the arguments
  u_NFW gives it, as the Taylor form sincos_cw is (both reach about 1.7:
  the reduction's and the summation's roundings);
* u_NFW over one common denominator per branch (P64) against the per-piece
  form (P64R, the Si/Ci quotients of src/generalFunctions.cu:44-87 taken one
  by one): equal to a few units in the last place of the result's
  cancellation scale, in each of the three branches (a1 >= 4, a1 < 4 <= a2,
  a2 < 4);
* both against the truncated-NFW transform with exact Si and Ci
  (src/generalFunctions.cu:6-30 at 40 digits): the same error, which is the
  rational approximants' own."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

mpmath = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("unfw") / "unfw_probe")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-x", "hip",
                        "--offload-arch=gfx950", "-Wno-unused-command-line-argument",
                        os.path.join(ROOT, "tests", "cxx", "unfw_probe.cpp"), "-o", exe],
                       capture_output=True, text=True)
    if r.returncode:
        pytest.skip("hipcc host build unavailable: " + r.stderr[-300:])
    return exe


def _hex(x):
    return f"{struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016x}"


def _run(exe, lines):
    out = subprocess.run([exe], input="\n".join(lines), capture_output=True, text=True,
                         check=True).stdout
    return [[struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0] for h in row.split()]
            for row in out.strip().split("\n")]


def _ulps(y, t):
    return float(abs(mpmath.mpf(y) - t) / mpmath.mpf(np.spacing(abs(float(t)))))


def test_sincos_mm_within_two_ulps(probe):
    mpmath.mp.dps = 40
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(1e-4, 4.0, 1500),            # u_NFW's a1 and a1 c below 4
                         np.exp(rng.uniform(np.log(4), np.log(300), 500)),
                         np.pi / 4 * np.arange(1, 41) * (1 + 1e-15)])   # reduction edges
    rows = _run(probe, [f"s {_hex(x)}" for x in xs])
    worst = np.zeros(4)
    for x, row in zip(xs, rows):
        ts, tc = mpmath.sin(mpmath.mpf(x)), mpmath.cos(mpmath.mpf(x))
        worst = np.maximum(worst, [_ulps(row[0], ts), _ulps(row[1], tc),
                                   _ulps(row[2], ts), _ulps(row[3], tc)])
    assert worst.max() < 2.0, worst


def _u_exact(a1, c):
    a1, c = mpmath.mpf(a1), mpmath.mpf(c)
    a2 = a1 * (1 + c)
    F = mpmath.log(1 + c) - c / (1 + c)
    return (mpmath.sin(a1) * (mpmath.si(a2) - mpmath.si(a1))
            + mpmath.cos(a1) * (mpmath.ci(a2) - mpmath.ci(a1)) - mpmath.sin(a1 * c) / a2) / F


def test_one_quotient_form(probe):
    mpmath.mp.dps = 40
    rng = np.random.default_rng(9)
    a1 = 10.0 ** rng.uniform(-4, 2, 1500)
    c = 10.0 ** rng.uniform(0, 1.7, 1500)
    rows = _run(probe, [f"u {_hex(x)} {_hex(y)}" for x, y in zip(a1, c)])
    seen = set()
    for x, y, (u, u0) in zip(a1, c, rows):
        b = 0 if x >= 4 else (1 if x * (1 + y) >= 4 else 2)
        seen.add(b)
        t = _u_exact(x, y)
        e, e0 = abs((u - t) / t), abs((u0 - t) / t)
        # the two forms agree far inside the approximants' error ...
        assert abs(u - u0) <= 2e-14 * abs(u0) + 1e-16, (x, y, u, u0)
        # ... and the one-quotient form is as close to the exact transform
        assert e <= max(2 * e0, 1e-14), (x, y, e, e0)
        # the reference's approximants: ~1e-14 outside the middle branch,
        # ~3e-11 inside it
        assert e < (1e-10 if b == 1 else 2e-14), (b, x, y, e)
    assert seen == {0, 1, 2}
