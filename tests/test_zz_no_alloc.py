"""g3l_eval_batch_noalloc, the allocation-free batched entry point, against
g3l_eval_batch through the C ABI: two calls in a row on one context, the
second with a smaller input, in both precisions; each result equals
g3l_eval_batch's for the same requests bit for bit (the same kernels refine
the same regions).  A workload whose buffers exceed the fixed scratch array
is rejected with G3L_ETOOBIG before any device work."""
import pytest

from conftest import ARCMIN, T, say

pytestmark = pytest.mark.gpu

FIELDS = ("value", "raw", "err", "neval", "rounds", "status")


def _requests(sid, trips, maxeval):
    from g3lhalo_amd import _lib
    return [_lib.make_request(stat, term, th * ARCMIN, hodset=sid, reltol=1e-4, maxeval=maxeval)
            for th in trips for stat in (0, 1) for term in (1, 2, 3)]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for f in FIELDS:
            assert getattr(x, f) == getattr(y, f), (f, getattr(x, f), getattr(y, f))


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_eval_batch(tables, precision):
    from g3lhalo_amd import _lib
    trips = T.triplets_550()[::61]          # 10 triplets x 6 kinds = 60 integrals
    c = _lib.Context(tables)
    try:
        sid = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        big = _requests(sid, trips, 300000)
        small = _requests(sid, trips[:3], 100000)
        say(f"noalloc: {len(big)} then {len(small)} integrals, precision {precision}")
        a_big = c.eval_batch_noalloc(big, precision)
        a_small = c.eval_batch_noalloc(small, precision)
        say("noalloc: the same requests through g3l_eval_batch")
        b_big = c.eval_batch(big, precision)
        b_small = c.eval_batch(small, precision)
    finally:
        c.close()
    _same(a_big, b_big)
    _same(a_small, b_small)
    assert all(r.neval > 0 for r in a_big + a_small)


def test_noalloc_rejects_workloads_beyond_the_array(tables):
    from g3lhalo_amd import _lib
    c = _lib.Context(tables)
    try:
        sid = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        # 40 5-D integrals at maxEval 5e6: 26,883 slots each, about 160 MB
        reqs = [_lib.make_request(0, 1, th * ARCMIN, hodset=sid, reltol=1e-4, maxeval=5000000)
                for th in T.triplets_550()[:40]]
        say("noalloc: oversized workload (rejected before device work)")
        with pytest.raises(_lib.G3LError) as e:
            c.eval_batch_noalloc(reqs)
        assert e.value.rc == _lib.G3L_ETOOBIG
        # the context and the array stay usable
        one = _lib.make_request(0, 1, T.triplets_550()[0] * ARCMIN, hodset=sid, reltol=1e-3,
                                maxeval=20000)
        r = c.eval_batch_noalloc([one])
        assert r[0].neval > 0 and r[0].neval == c.eval_batch([one])[0].neval
    finally:
        c.close()
