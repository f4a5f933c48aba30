"""Pins the oracle's cubature driver (restated from src/hcubature.cpp).

* Bit-identical value, error, evaluation count and batch-size trace against
  the reference's own integrator compiled from src/hcubature.cpp
  (oracle/_ref), on Genz test families and on every physics integrand.
* Closed-form integrals of the Genz families.
* Exactness of the Genz-Malik 7/5 rule on monomials and of Gauss-Kronrod 15.
"""
import math
import os
import sys

import numpy as np
import pytest

from conftest import ARCMIN, T

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle"))
import oracle as O  # noqa: E402

REF = O.reference_hcubature()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (no reference sources)")


def genz_exact(family, a, u):
    """Closed forms over [0,1]^d."""
    a = np.asarray(a, float)
    u = np.asarray(u, float)
    if family == 1:  # product peak
        return float(np.prod(a * (np.arctan(a * (1 - u)) + np.arctan(a * u))))
    if family == 3:  # Gaussian
        return float(np.prod(math.sqrt(math.pi) / (2 * a) *
                             (np.array([math.erf(x) for x in a * (1 - u)]) +
                              np.array([math.erf(x) for x in a * u]))))
    if family == 0:  # oscillatory: Re prod over dims of int e^{i a x} times e^{i 2 pi u0}
        z = complex(math.cos(2 * math.pi * u[0]), math.sin(2 * math.pi * u[0]))
        for aj in a:
            z *= (complex(math.cos(aj), math.sin(aj)) - 1) / complex(0, aj)
        return z.real
    if family == 4:  # C0
        return float(np.prod((2 - np.exp(-a * u) - np.exp(-a * (1 - u))) / a))
    raise ValueError(family)


def test_points_per_region():
    for d, n in [(1, 15), (2, 17), (3, 33), (5, 93), (6, 149), (7, 241)]:
        assert O.lib().orc_points_per_region(d) == n
        if d > 1:
            assert n == 1 + 4 * d + 2 * d * (d - 1) + 2 ** d


@pytest.mark.parametrize("family", [0, 1, 3, 4])
@pytest.mark.parametrize("dim", [2, 3, 5])
def test_genz_closed_form(family, dim):
    rng = np.random.default_rng(100 * family + dim)
    a = rng.uniform(1.0, 4.0, dim)
    u = rng.uniform(0.1, 0.9, dim)
    lo, hi = np.zeros(dim), np.ones(dim)
    val, err, ne, nc = O.genz(family, a, u, lo, hi, maxeval=2_000_000, reltol=1e-7)
    exact = genz_exact(family, a, u)
    assert abs(val - exact) <= max(2 * err, 1e-12 * abs(exact)), (val, exact, err)
    if err <= 1e-7 * abs(val):  # converged within the budget
        assert abs(val - exact) <= 1e-6 * abs(exact)


def test_gk15_1d_closed_form():
    val, err, ne, nc = O.genz(0, [5.0], [0.0], [0.0], [3.0], maxeval=0, reltol=1e-10)
    exact = math.sin(15.0) / 5.0  # int_0^3 cos(5 x) dx
    assert abs(val - exact) < 1e-9


@pytest.mark.parametrize("dim", [2, 3, 5, 6, 7])
def test_gm_rule_exact_on_monomials(dim):
    """One region (maxeval = one rule application): degree <= 7 integrated
    exactly by the 7th-degree rule; degree <= 5 gives a zero error estimate."""
    npt = O.lib().orc_points_per_region(dim)
    rng = np.random.default_rng(dim)
    lo = rng.uniform(-1, 0, dim)
    hi = lo + rng.uniform(0.5, 2.0, dim)
    for deg in range(0, 8):
        e = np.zeros(dim)
        for _ in range(deg):
            e[rng.integers(dim)] += 1
        val, err, ne, nc = O.genz(5, e, np.zeros(dim), lo, hi, maxeval=npt, reltol=0.0)
        assert ne == npt
        exact = float(np.prod((hi ** (e + 1) - lo ** (e + 1)) / (e + 1)))
        scale = float(np.prod(np.maximum(np.abs(hi), np.abs(lo)) ** e * (hi - lo)))
        assert abs(val - exact) <= 1e-13 * scale, (deg, e, val, exact)
        if deg <= 5:
            assert err <= 1e-13 * scale, (deg, err)
    # degree 8: the 5th/7th-degree difference is non-zero
    e = np.zeros(dim)
    e[0] = 8
    val, err, ne, nc = O.genz(5, e, np.zeros(dim), lo, hi, maxeval=npt, reltol=0.0)
    assert err > 0


def test_gk15_exactness():
    # Kronrod 15 exact to degree 22, Gauss 7 to degree 13: err at rounding level
    for deg in (0, 5, 13):
        val, err, ne, nc = O.genz(5, [float(deg)], [0.0], [0.2], [1.7], maxeval=15, reltol=0.0)
        exact = (1.7 ** (deg + 1) - 0.2 ** (deg + 1)) / (deg + 1)
        assert abs(val - exact) < 1e-14 * max(1, exact)
        assert err < 1e-12 * max(1, exact)
    val, err, ne, nc = O.genz(5, [22.0], [0.0], [0.2], [1.7], maxeval=15, reltol=0.0)
    assert abs(val - (1.7 ** 23 - 0.2 ** 23) / 23) < 1e-12 * 1.7 ** 23
    val, err, ne, nc = O.genz(5, [20.0], [0.0], [0.2], [1.7], maxeval=15, reltol=0.0)
    assert err > 1e-10 * 1.7 ** 20  # Gauss 7 no longer exact


@needs_ref
@pytest.mark.parametrize("family", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dim,maxeval,tol", [(1, 0, 1e-9), (2, 50000, 1e-6), (4, 200000, 1e-5),
                                             (5, 300000, 1e-4), (7, 300000, 1e-4)])
def test_driver_bitwise_vs_reference_genz(family, dim, maxeval, tol):
    rng = np.random.default_rng(family * 31 + dim)
    a = rng.uniform(2.0, 9.0, dim)
    u = rng.uniform(0.0, 1.0, dim)
    lo, hi = -rng.uniform(0, 1, dim), rng.uniform(1, 2, dim)
    mine = O.genz(family, a, u, lo, hi, maxeval=maxeval, reltol=tol)
    ref = O.genz(family, a, u, lo, hi, maxeval=maxeval, reltol=tol, hcub=REF)
    # the corner peak has a pole inside these boxes: inf/NaN must propagate alike
    assert np.array_equal(np.array(mine, float), np.array(ref, float), equal_nan=True)


@needs_ref
@pytest.mark.parametrize("stat,term", [(0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3)])
@pytest.mark.parametrize("types", [(1, 2), (1, 1), (2, 2)])
def test_driver_bitwise_vs_reference_terms(tables, oracle_model, stat, term, types):
    if stat == 1 and types != (1, 2):
        pytest.skip("NMapMap has one HOD")
    th = np.array([0.7, 3.0, 11.0]) * ARCMIN
    a = oracle_model.eval_term(stat, term, th, *types, reltol=1e-4, maxeval=150000,
                               trace_cap=200)
    b = oracle_model.eval_term(stat, term, th, *types, reltol=1e-4, maxeval=150000,
                               trace_cap=200, hcub=REF)
    assert a["raw"] == b["raw"] and a["err"] == b["err"]
    assert a["neval"] == b["neval"] and a["ncalls"] == b["ncalls"]
    assert np.array_equal(a["trace"], b["trace"])


@needs_ref
def test_densities_bitwise_vs_reference(tables, oracle_model):
    m = O.Model(tables, hcub=REF)
    m.set_hods(T.FIDUCIAL_HOD, (0.6, 0.8, 3e11, 0.3, 2e13, 0.8))
    mine = O.Model(tables)
    mine.set_hods(T.FIDUCIAL_HOD, (0.6, 0.8, 3e11, 0.3, 2e13, 0.8))
    for x, y in zip(m.densities(), mine.densities()):
        assert np.array_equal(x, y)


def test_maxeval_accounting(oracle_model):
    """numEval counts the root, then 2 x points per popped region, and the
    last batch stops at the first pop that reaches the budget."""
    th = np.array([1.0, 1.0, 1.0]) * ARCMIN
    for me in (93, 94, 1000, 20000):
        r = oracle_model.eval_term(0, 1, th, reltol=1e-9, maxeval=me, trace_cap=100)
        ne = r["neval"]
        assert (ne - 93) % 186 == 0
        assert ne >= me or ne == 93
        assert ne - 186 < me or me <= 93
        assert sum(r["trace"]) == ne
