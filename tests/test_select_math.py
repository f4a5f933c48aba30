"""Host check of the batch selection's fixed-point error sums
(g3lhalo_amd/csrc/select_math.h, used by k_select): no wrap-around for up to
2^26 equal error keys (an integral's largest slot capacity), and the pop
count derived from the fixed-point total equals the count of the
reference's pop loop (src/hcubature.cpp:971-985) for equal errors."""
import os
import subprocess

from conftest import ROOT


def test_fixed_point_selection_sums(tmp_path):
    exe = str(tmp_path / "select_probe")
    r = subprocess.run(["g++", "-std=c++20", "-O1", os.path.join(ROOT, "tests", "cxx",
                                                                  "select_probe.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
