"""The NNMap 3-halo parity bar (tests/ulp_bar.py) on planted deviations:
it accepts a device value that lands where one of the oracle's one-ulp runs
lands, or twice as far, and rejects one beyond that, an extra integral
beyond 1e-4, or a deviation larger than the oracle's own largest."""
import numpy as np
import pytest

from ulp_bar import check_three_halo


def _case():
    rng = np.random.default_rng(7)
    o = rng.uniform(1.0, 2.0, 40)
    ulp = np.tile(o, (4, 1)) * (1 + 1e-15)
    ulp[0, 3] = o[3] * (1 + 3e-4)       # the oracle's one-ulp runs move integrals 3 and 9
    ulp[0, 9] = o[9] * (1 - 5e-4)
    ulp[2, 9] = o[9] * (1 + 1e-4 / 2)
    return o, ulp


def test_accepts_devices_within_the_oracles_own_spread():
    o, ulp = _case()
    dev = o * (1 + 1e-14)
    dev[3] = o[3] * (1 - 3e-4)          # the other side: 2x the oracle's move from its run
    dev[9] = o[9] * (1 - 5e-4)          # where the oracle's own run landed
    d = check_three_halo(dev, o, ulp, "planted")
    assert d[3] == pytest.approx(3e-4) and d[9] == pytest.approx(5e-4)


@pytest.mark.parametrize("i, rel", [(3, 7e-4),    # beyond twice the oracle's move
                                    (5, 2e-4)])   # an integral the oracle does not move
def test_rejects_deviation_beyond_the_bar(i, rel):
    o, ulp = _case()
    dev = o.copy()
    dev[i] = o[i] * (1 + rel)
    with pytest.raises(AssertionError):
        check_three_halo(dev, o, ulp, "planted")


def test_rejects_more_movers_than_one_oracle_run():
    o, ulp = _case()
    ulp[1, 5] = o[5] * (1 + 2e-4)       # integral 5 moves, but in another seed's run
    dev = o.copy()
    dev[3], dev[9], dev[5] = o[3] * (1 + 3e-4), o[9] * (1 - 5e-4), o[5] * (1 + 2e-4)
    with pytest.raises(AssertionError):
        check_three_halo(dev, o, ulp, "planted")
