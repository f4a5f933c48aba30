"""g3l::log_pos, the ln x of Ci(x) inside u_NFW (fdlibm's e_log.c method),
built for the host from the same source and checked against mpmath at 40
digits: within one unit in the last place over the arguments u_NFW gives it,
(0, 4), and beyond; exact at 1 and at powers of two times ln 2's split."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

mpmath = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("logp") / "log_probe")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-std=c++17", "-x", "hip",
                        "--offload-arch=gfx950", "-Wno-unused-command-line-argument",
                        os.path.join(ROOT, "tests", "cxx", "log_probe.cpp"), "-o", exe],
                       capture_output=True, text=True)
    if r.returncode:
        pytest.skip("hipcc host build unavailable: " + r.stderr[-300:])
    return exe


def _run(exe, xs):
    inp = "\n".join(f"{struct.unpack('<Q', struct.pack('<d', float(x)))[0]:016x}" for x in xs)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    return [struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0] for h in out.split()]


def _ulp_err(x, y):
    mpmath.mp.dps = 40
    t = mpmath.log(mpmath.mpf(x))
    return float(abs(mpmath.mpf(y) - t) / mpmath.mpf(np.spacing(abs(float(t)))))


def test_within_one_ulp(probe):
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(1e-3, 4.0, 3000),                 # u_NFW's range
                         np.exp(rng.uniform(-30, 30, 1000)),           # beyond
                         np.sqrt(0.5) * (1 + rng.uniform(-1e-12, 1e-12, 200)),  # reduction edge
                         1 + rng.uniform(-1e-6, 1e-6, 200)])           # near log = 0
    ys = _run(probe, xs)
    err = max(_ulp_err(x, y) for x, y in zip(xs, ys) if x != 1.0)
    assert err < 1.0, err


def test_special_points(probe):
    ys = _run(probe, [1.0, 2.0, 0.5, 4.0])
    assert ys[0] == 0.0
    # k ln2 in two parts: within an ulp of the correctly rounded k ln 2
    for k, y in zip((1, -1, 2), ys[1:]):
        assert _ulp_err(2.0 ** k, y) < 1.0
