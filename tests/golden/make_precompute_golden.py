"""Writes tests/golden/precompute_<name>/: the inputs and the output files of
the reference's table precomputation (doPrecomputations.x, built from the
reference's own sources into oracle/_ref by `make -C oracle ref`) for two
small configurations -- the fiducial flat LCDM cosmology at 16 bins, the
reference's example cosmology (scripts_examples/cosmo.param values) at 8
bins, and the fiducial cosmology on the grid that tests/test_gpu_pipeline.py
hands to the drop-in drivers (redshifts 0.01-2.01, whose last abscissa prints
above its exact value, so src/Function.cpp:55-57 reads it back).  The device precomputation (g3l_precompute) is checked against these
files.  A few seconds on 4 cores.

Usage: python tests/golden/make_precompute_golden.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "doPrecomputations.x")

OUTPUTS = ("g.dat", "w.dat", "dwdz.dat", "hmf.dat", "pk.dat", "bh.dat", "conc.dat")

CONFIGS = {
    "fiducial16": {"cosmo": {"Omega_m": 0.3, "Omega_Lambda": 0.7, "Omega_b": 0.045,
                             "sigma8": 0.8, "H0": 70, "ns": 0.96},
                   "zmin": 0.001, "zmax": 2.0, "mmin": 1e10, "mmax": 1e17, "kmin": 1e-4,
                   "kmax": 1e4, "nbins": 16},
    "example8": {"cosmo": {"Omega_m": 0.25, "Omega_Lambda": 0.75, "Omega_b": 0.045,
                           "sigma8": 0.9, "H0": 73, "ns": 1.0},
                 "zmin": 0.1, "zmax": 2.0, "mmin": 1e10, "mmax": 1e17, "kmin": 0.01,
                 "kmax": 1000.0, "nbins": 8},
    "pipeline16": {"cosmo": {"Omega_m": 0.3, "Omega_Lambda": 0.7, "Omega_b": 0.045,
                             "sigma8": 0.8, "H0": 70, "ns": 0.96},
                   "zmin": 0.01, "zmax": 2.01, "mmin": 1e10, "mmax": 1e17, "kmin": 1e-4,
                   "kmax": 1e4, "nbins": 16},
}


def source_nz(path):
    """n(z) proportional to z^2 exp(-(z/0.5)^1.5) on 100 points in [0.01, 3],
    normalised on its own grid."""
    z = np.linspace(0.01, 3.0, 100)
    n = z ** 2 * np.exp(-(z / 0.5) ** 1.5)
    n /= n.sum() * (z[-1] - z[0]) / len(z)
    with open(path, "w") as f:
        for a, b in zip(z, n):
            f.write(f"{a:.17g} {b:.17g}\n")


def main():
    if not os.path.exists(REF):
        sys.exit("oracle/_ref/doPrecomputations.x missing: run `make -C oracle ref` with the "
                 "reference sources present")
    for name, cfg in CONFIGS.items():
        d = os.path.join(HERE, f"precompute_{name}")
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "cosmo.param"), "w") as f:
            for k, v in cfg["cosmo"].items():
                f.write(f"{k} {v}\n")
        source_nz(os.path.join(d, "nz_sources.dat"))
        with open(os.path.join(d, "config.json"), "w") as f:
            json.dump(cfg, f, indent=1)
        args = [REF, "cosmo.param", "nz_sources.dat", repr(cfg["zmin"]), repr(cfg["zmax"]),
                repr(cfg["mmin"]), repr(cfg["mmax"]), repr(cfg["kmin"]), repr(cfg["kmax"]),
                str(cfg["nbins"]), *OUTPUTS]
        subprocess.run(args, cwd=d, check=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
        print("wrote", d)


if __name__ == "__main__":
    main()
