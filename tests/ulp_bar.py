"""The parity bar of the NNMap 3-halo integrals (shared by the full-size and
HOD-sweep tests).

Every NNMap 3-halo integral of these workloads stops at the evaluation cap
unconverged, and its refinement is chaotic at the last place: a one-ulp
change in any integrand factor can change which of two near-tied regions is
split, after which two runs drift apart by up to the integral's own error.
The oracle shows how far: recomputed with every exp, log, pow, sin, cos and
erf result moved by at most one unit in the last place (its sensitivity
probe, several seeds), it moves most of these integrals, some by more than
1e-4.  The device is one more implementation whose math library differs
from the oracle's in last places, so on such an integral it can land as far
from the oracle as two one-ulp runs land from each other: up to the sum of
their moves.  The moves come in near-equal steps (one flipped split each),
and one run can take two where another takes one: on one sweep integral the
oracle's seeds move it by 2.965e-3, 2.966e-3 and 5.916e-3.

The bar, per integral: within 1e-4 of the oracle, or within twice the
largest move the oracle's own one-ulp runs make on that same integral (to
1 %, the variation of one step between runs); and,
checked by the caller, within 10 % of the oracle's error estimate.  In
aggregate: no more integrals beyond 1e-4, and none further, than one of the
oracle's own one-ulp runs.
"""
import numpy as np


def check_three_halo(dev, oracle, ulp, label):
    """dev, oracle: the 3-halo values; ulp: the oracle's values under its
    one-ulp probe, one row per seed.  Prints the comparison, asserts the bar."""
    dev, oracle, ulp = np.asarray(dev), np.asarray(oracle), np.asarray(ulp)
    assert ulp.ndim == 2 and ulp.shape[0] >= 4 and ulp.shape[1] == len(oracle)
    d = np.abs(dev - oracle) / np.abs(oracle)
    du = np.abs(ulp - oracle) / np.abs(oracle)
    spread = du.max(axis=0)
    out = np.flatnonzero(d > 1e-4)
    n_seed = int((du > 1e-4).sum(axis=1).max())
    print(f"{label}: {len(out)} of {len(d)} above 1e-4 (max {d.max():.2e}, median "
          f"{np.median(d):.2e}); the oracle's one-ulp runs: at most {n_seed} per seed, "
          f"{int((spread > 1e-4).sum())} distinct, max {spread.max():.2e}")
    for i in out:
        print(f"  integral {i}: rel {d[i]:.2e}, the oracle's largest one-ulp move "
              f"{spread[i]:.2e} (ratio {d[i] / spread[i]:.2f})")
    beyond = out[d[out] > 2.02 * spread[out]]
    assert len(beyond) == 0, beyond
    assert len(out) <= n_seed and d.max() <= spread.max()
    return d
