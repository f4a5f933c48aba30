"""Writes tests/golden/cubature_ref.json: outputs of the reference's own
adaptive cubature (src/hcubature.cpp, compiled by oracle/Makefile into
oracle/_ref/libref_hcubature.so) on Genz test families and on the path's
integrands, so that the oracle can be checked against the reference where
the reference sources are absent.

Each case records val, err, the evaluation count, the number of integrand
calls and the batch-size trace (points per call), with floats as exact hex
strings.  The physics integrands are the oracle's (the reference's integrand
sources need CUDA headers this image lacks), so these vectors pin the
integration driver: region rule, splitting, heap order, Gladwell batching
and the evaluation budget.

Usage (needs /root/reference): make -C oracle all ref && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import oracle as O  # noqa: E402
from g3lhalo_amd import tables as T  # noqa: E402

GENZ = [  # family, dim, maxeval, reltol, seed
    (0, 3, 60000, 1e-7, 1), (1, 5, 200000, 1e-5, 2), (2, 4, 100000, 1e-6, 3),
    (3, 6, 300000, 1e-4, 4), (4, 7, 250000, 1e-4, 5), (3, 1, 0, 1e-10, 6),
    (1, 2, 0, 1e-9, 7),
]
HOD2 = (0.6, 0.8, 3e11, 0.3, 2e13, 0.8)
TERMS = [  # stat, term, theta [arcmin], types, reltol, maxeval
    (0, 1, (1.0, 1.0, 1.0), (1, 2), 1e-3, 40000),
    (0, 1, (0.7, 3.0, 11.0), (1, 1), 1e-4, 60000),
    (0, 2, (2.0, 5.0, 5.0), (1, 2), 1e-4, 60000),
    (0, 3, (4.0, 4.0, 9.0), (2, 2), 1e-4, 60000),
    (1, 1, (1.5, 2.5, 8.0), (1, 2), 1e-4, 60000),
    (1, 2, (3.0, 3.0, 3.0), (1, 2), 1e-4, 60000),
    (1, 3, (6.0, 2.0, 4.0), (1, 2), 1e-4, 60000),
]
TRACE_CAP = 4096


def hx(v):
    return float(v).hex()


def main():
    O.build()
    ref = O.reference_hcubature()
    if ref is None:
        sys.exit("oracle/_ref/libref_hcubature.so is not built (needs the reference sources)")
    out = {"generator": "tests/golden/make_golden.py", "genz": [], "terms": [], "densities": {}}
    for fam, dim, maxeval, tol, seed in GENZ:
        rng = np.random.default_rng(seed)
        a = rng.uniform(2.0, 9.0, dim)
        u = rng.uniform(0.0, 1.0, dim)
        lo, hi = -rng.uniform(0, 1, dim), rng.uniform(1, 2, dim)
        val, err, ne, nc = O.genz(fam, a, u, lo, hi, maxeval=maxeval, reltol=tol, hcub=ref)
        out["genz"].append({"family": fam, "dim": dim, "a": [hx(x) for x in a],
                            "u": [hx(x) for x in u], "lo": [hx(x) for x in lo],
                            "hi": [hx(x) for x in hi], "maxeval": maxeval, "reltol": tol,
                            "val": hx(val), "err": hx(err), "neval": int(ne), "ncalls": int(nc)})
    t = T.make_tables()
    out["tables_checksum"] = T.checksum(t)
    two = O.Model(t, hcub=ref)
    two.set_hods(T.FIDUCIAL_HOD, HOD2)
    same = O.Model(t, hcub=ref)
    same.set_hods(T.FIDUCIAL_HOD, identical=True)
    for stat, term, th, types, tol, maxeval in TERMS:
        m = same if stat == 1 else two
        r = m.eval_term(stat, term, np.array(th) * T.ARCMIN_IN_RAD, *types, reltol=tol,
                        maxeval=maxeval, trace_cap=TRACE_CAP, hcub=ref)
        assert len(r["trace"]) < TRACE_CAP
        out["terms"].append({"stat": stat, "term": term, "theta_arcmin": list(th),
                             "types": list(types), "reltol": tol, "maxeval": maxeval,
                             "hod2": "HOD2" if stat == 0 else "same",
                             "value": hx(r["value"]), "raw": hx(r["raw"]), "err": hx(r["err"]),
                             "neval": r["neval"], "ncalls": r["ncalls"],
                             "trace": [int(x) for x in r["trace"]]})
    rho, n1, n2 = two.densities()
    out["densities"] = {"hod1": "FIDUCIAL", "hod2": list(HOD2), "rho_bar": [hx(x) for x in rho],
                        "n_bar1": [hx(x) for x in n1], "n_bar2": [hx(x) for x in n2]}
    out["hod2"] = list(HOD2)
    path = os.path.join(HERE, "cubature_ref.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
