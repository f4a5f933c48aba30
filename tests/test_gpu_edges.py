"""Edge cases of the device path against the oracle, through the C ABI.

* Evaluation budgets below one rule: the reference evaluates the root region,
  counts its points and never enters the refinement loop
  (src/hcubature.cpp:931-938), so the result is the root region's estimate
  after 93/149/241 evaluations.
* Budgets that end one region into the first batch, and inside a later batch:
  the Gladwell batch stops when the budget is spent
  (src/hcubature.cpp:971-985), so the evaluation counts must match exactly.
* A zero comoving distance in the first redshift bin (the input a zero zmin
  gives the reference): k = l / w is infinite there and the integrand
  non-finite.  The convergence test of src/converged.h:11-13 holds for a NaN
  error total, so the integral stops in the round that first produces one,
  with a non-finite value; the device reports G3L_NONFINITE, stops after the
  oracle's evaluation count and reads no table entry outside the tables
  (every index is clamped).  The tables have 64 bins: at 32 the root
  region's centre mass, log10 m = 10 + 3.390625, sits on an exact rounding
  tie of the mass index (3.390625 x 32 / 7 = 15.5), which the last bit of
  pow(10, x) decides; the device's pow and the host library's differ in that
  bit there, so the two take neighbouring bins and refine differently from
  the first split on.
* An empty batch.
"""
import numpy as np
import pytest

from conftest import ARCMIN, T, say

pytestmark = pytest.mark.gpu

TERMS = [(0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3)]
PTS = {1: 93, 2: 149, 3: 241}      # Genz-Malik points per 5/6/7-D region
THETA = np.array([2.0, 4.0, 8.0]) * ARCMIN
G3L_MAXEVAL_HIT, G3L_NONFINITE = 5, 6


@pytest.fixture(scope="module")
def ctx(tables):
    from g3lhalo_amd import _lib
    say("edges: context + fiducial HOD set")
    c = _lib.Context(tables)
    assert c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD) == 0
    yield c
    c.close()


def _budgets(term):
    p = PTS[term]
    # below one rule; exactly one rule; one region cut; five full batches of
    # regions plus a few points (ends inside the next batch)
    return (1, p, p + 1, p + 10 * p + 7)


@pytest.mark.parametrize("stat,term", TERMS)
def test_small_budgets_match_oracle(ctx, oracle_model, stat, term):
    from g3lhalo_amd import _lib
    budgets = _budgets(term)
    reqs = [_lib.make_request(stat, term, THETA, reltol=1e-6, maxeval=b) for b in budgets]
    say(f"edges: budgets {budgets} stat={stat} term={term}")
    res = ctx.eval_batch(reqs)
    for b, r in zip(budgets, res):
        o = oracle_model.eval_term(stat, term, THETA, reltol=1e-6, maxeval=b)
        d = abs(r.value - o["value"]) / abs(o["value"])
        print(f"maxeval {b}: neval {r.neval} oracle {o['neval']}, rel {d:.2e}, status {r.status}")
        assert r.neval == o["neval"]
        assert r.status == G3L_MAXEVAL_HIT
        assert d < 1e-9
        if b <= PTS[term]:
            assert r.neval == PTS[term] and r.rounds == 1
    # the plain form on the smallest and the largest budget
    for b in (budgets[0], budgets[-1]):
        p = ctx.eval_plain(_lib.make_request(stat, term, THETA, reltol=1e-6, maxeval=b))
        q = [r for bb, r in zip(budgets, res) if bb == b][0]
        assert p.neval == q.neval
        assert abs(p.value - q.value) <= 1e-12 * abs(q.value)


@pytest.fixture(scope="module")
def zero_w():
    import dataclasses
    t = T.make_tables(nbins=64)
    w = t.w.copy()
    w[0] = 0.0
    return dataclasses.replace(t, w=w)


@pytest.mark.parametrize("stat,term", [(0, 1), (0, 2), (0, 3), (1, 1)])
def test_zero_comoving_distance_is_nonfinite(zero_w, stat, term):
    import oracle as O
    from g3lhalo_amd import _lib
    m = O.Model(zero_w)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    o = m.eval_term(stat, term, THETA, reltol=1e-3, maxeval=2000000)
    assert not np.isfinite(o["value"])
    say(f"edges: zero comoving distance stat={stat} term={term}")
    c = _lib.Context(zero_w)
    try:
        assert c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD) == 0
        req = _lib.make_request(stat, term, THETA, reltol=1e-3, maxeval=2000000)
        r = c.eval_batch([req])[0]
        p = c.eval_plain(req)
    finally:
        c.close()
    print(f"neval device {r.neval} plain {p.neval} oracle {o['neval']}; status {r.status}")
    for x in (r, p):
        assert x.status == G3L_NONFINITE
        assert not np.isfinite(x.value)
        assert x.neval == o["neval"]


def test_empty_batch(ctx):
    from g3lhalo_amd import _lib
    say("edges: empty batch")
    assert ctx.eval_batch([]) == []
    # the context still works afterwards
    r = ctx.eval_batch([_lib.make_request(0, 1, THETA, reltol=1e-6, maxeval=1)])[0]
    assert r.neval == PTS[1]
