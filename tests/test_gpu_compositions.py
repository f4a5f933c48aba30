"""The device integrands' compositions (g3l_integrand, the C ABI's integrand
hook) against what the mathematics fixes (tests/compositions.py): exact
power-of-two degrees under doubling each input table, in both precision
policies (the products of table values are FP64 in both); the mass and
redshift index of each look-up from local versions of those scalings; the
NMapMap 3-halo mass mix (Q4); the HOD limit (NNMap invariant, NMapMap
doubled).  The device's n_bar(z) (batched Gauss-Kronrod on the device) must
double exactly with hmf for the hmf degrees to come out."""
import pytest

import compositions as K
from conftest import T, say

pytestmark = pytest.mark.gpu

_open = []


@pytest.fixture(scope="module", autouse=True)
def _close_contexts():
    yield
    while _open:
        _open.pop().close()


def _evaluator(t, ha, hb, A, eps, precision):
    from g3lhalo_amd import _lib
    c = _lib.Context(t)
    _open.append(c)
    sid = c.hodset(ha, hb, A=A, eps=eps, identical=False)

    def f(stat, term, x):
        req = _lib.make_request(stat, term, K.THETA, hodset=sid, type1=1, type2=2)
        return c.integrand(req, x, precision)
    return f


def maker(precision):
    return lambda t: _evaluator(t, K.HOD_A, K.HOD_B, K.A_CORR, K.EPS_CORR, precision)


@pytest.fixture(scope="module")
def t():
    return T.make_tables()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", sorted(K.DEGREE))
def test_global_table_scaling(t, name, precision):
    say(f"compositions: {name} doubled, precision {precision}")
    K.check_global_scaling(maker(precision), t, name)


def test_bh_mass_column(t):
    say("compositions: b_h column")
    K.check_bh_column(maker(0), t, 60)


def test_concentration_mass_column(t):
    say("compositions: concentration column")
    K.check_conc_column(maker(0), t, 61)


@pytest.mark.parametrize("name", ["hmf", "P_lin"])
def test_redshift_row(t, name):
    say(f"compositions: {name} row")
    K.check_row(maker(0), t, name, 70)


def test_hod_limit(t):
    say("compositions: HOD limit")
    K.check_hod_limit(lambda tt, ha, hb: _evaluator(tt, ha, hb, 0.0, 0.0, 0), t, rtol=1e-9)
