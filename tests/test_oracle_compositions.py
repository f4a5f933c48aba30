"""The oracle's six integrand compositions against what the mathematics fixes
(tests/compositions.py): exact power-of-two degrees under doubling each input
table, the mass and redshift indices of each look-up from local versions of
those scalings, the NMapMap 3-halo mass mix (Q4) from a concentration change
in one mass column, and the HOD limit in which every NNMap term is invariant
and every NMapMap term doubles.  These pin the compositions without
restating them: a dropped or doubled factor, a wrong mass slot or index, or a
lost second n_bar division changes an exponent."""
import pytest

import compositions as K
from conftest import T

import oracle as O


def make(t):
    m = O.Model(t)
    m.set_hods(K.HOD_A, K.HOD_B, A=K.A_CORR, eps=K.EPS_CORR, identical=False)
    return lambda stat, term, x: m.integrand(stat, term, K.THETA, x, 1, 2)


def make_hod(t, ha, hb):
    m = O.Model(t)
    m.set_hods(ha, hb, identical=False)
    return lambda stat, term, x: m.integrand(stat, term, K.THETA, x, 1, 2)


@pytest.fixture(scope="module")
def t():
    return T.make_tables()


@pytest.mark.parametrize("name", sorted(K.DEGREE))
def test_global_table_scaling(t, name):
    K.check_global_scaling(make, t, name)


def test_bh_mass_column(t):
    K.check_bh_column(make, t, 60)


def test_concentration_mass_column(t):
    K.check_conc_column(make, t, 61)


@pytest.mark.parametrize("name", ["hmf", "P_lin"])
def test_redshift_row(t, name):
    K.check_row(make, t, name, 70)


def test_hod_limit(t):
    K.check_hod_limit(make_hod, t, rtol=0.0)


def test_checks_catch_a_planted_slip(t):
    """The degree check fails for an integrand that divides by n_bar once
    where NMapMap divides twice: a function that multiplies the oracle's
    NMapMap 1-halo value back by n_bar(z) (the slip the Q3 quirk register
    guards against) gets exponent 0 for hmf instead of -1."""
    import numpy as np

    def slipped(tt):
        m = O.Model(tt)
        m.set_hods(K.HOD_A, K.HOD_B, A=K.A_CORR, eps=K.EPS_CORR, identical=False)
        _, n1, _ = m.densities()

        def f(stat, term, x):
            v = m.integrand(stat, term, K.THETA, x, 1, 2)
            if stat == 1:
                zi = np.clip(np.round((x[:, -1] - tt.zmin) * tt.nbins / (tt.zmax - tt.zmin))
                             .astype(int), 0, tt.nbins - 1)
                v = v * n1[zi]
            return v
        return f

    with pytest.raises(AssertionError):
        K.check_global_scaling(slipped, t, "hmf")
