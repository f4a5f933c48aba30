"""Continuous batching (g3l_server_*): integrals submitted while others
refine are admitted as pool memory frees up, and every result equals the
result of the same integral in g3l_eval_batch -- bit for bit, since an
integral's refinement depends only on its own regions."""
import numpy as np
import pytest

from conftest import ARCMIN, T, say

pytestmark = pytest.mark.gpu

MAXEVAL = 300000


@pytest.fixture(scope="module")
def sctx(tables):
    from g3lhalo_amd import _lib
    say("server: context + two HOD sets")
    c = _lib.Context(tables)
    a = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    b = c.hodset(T.FIDUCIAL_HOD, tuple(T.hod_prior_draws(1, seed=11)[0]), A=0.3, eps=0.2)
    yield c, a, b
    c.close()


def _requests(a, b):
    from g3lhalo_amd import _lib
    th1 = np.array([2.0, 3.0, 5.0]) * ARCMIN
    th2 = np.array([6.0, 6.0, 1.5]) * ARCMIN
    r1 = [_lib.make_request(0, t, th, hodset=a, reltol=1e-4, maxeval=MAXEVAL)
          for th in (th1, th2) for t in (1, 2, 3)]
    r2 = [_lib.make_request(1, t, th1, hodset=a, reltol=1e-4, maxeval=MAXEVAL) for t in (1, 2, 3)]
    r3 = [_lib.make_request(0, t, th2, hodset=b, type1=t1, type2=t2, reltol=1e-4, maxeval=MAXEVAL)
          for t in (1, 2, 3) for (t1, t2) in ((1, 2), (1, 1), (2, 2))]
    return [r1, r2, r3]


def _same(got, ref):
    for g, r in zip(got, ref):
        assert g.status == r.status
        assert g.neval == r.neval
        assert g.value == r.value and g.err == r.err


@pytest.mark.parametrize("precision", [0, 1])
def test_server_results_equal_batch(sctx, precision):
    from g3lhalo_amd import _lib
    c, a, b = sctx
    groups = _requests(a, b)
    say(f"server precision={precision}: staggered submissions")
    c.set_mixed_policy(0x7e)   # every term in the FP32-bulk form under MIXED
    srv = _lib.Server(c, pool_slots=1 << 22, max_integrals=4096, precision=precision)
    try:
        srv.submit(groups[0], tag=1)
        done = {}
        for _ in range(3):
            srv.step()
        srv.submit(groups[1], tag=2)
        srv.step()
        srv.submit(groups[2], tag=3)
        for _ in range(10000):
            st = srv.step()
            done.update(dict(srv.poll()))
            if st.in_flight == 0 and st.pending == 0 and len(done) == 3:
                break
        assert sorted(done) == [1, 2, 3]
    finally:
        srv.close()
    for tag, reqs in zip((1, 2, 3), groups):
        _same(done[tag], c.eval_batch(reqs, precision=precision))


def test_server_admits_by_memory(sctx):
    from g3lhalo_amd import _lib
    c, a, b = sctx
    reqs = [r for g in _requests(a, b)[:2] for r in g]
    # room for two 5-D integrals' worth of slots: the rest wait their turn
    cap5 = 2 + (MAXEVAL + 93) // (2 * 93)
    say("server: small pool")
    srv = _lib.Server(c, pool_slots=2 * cap5, max_integrals=64)
    try:
        srv.submit(reqs, tag=7)
        peak, steps = 0, 0
        done = []
        while not done:
            st = srv.step()
            peak = max(peak, st.in_flight)
            steps += 1
            done = srv.poll()
            assert steps < 20000
        assert 1 <= peak < len(reqs)
        assert done[0][0] == 7
    finally:
        srv.close()
    _same(done[0][1], c.eval_batch(reqs))


def test_server_rejects_integral_larger_than_pool(sctx):
    from g3lhalo_amd import _lib
    c, a, _ = sctx
    srv = _lib.Server(c, pool_slots=1000, max_integrals=16)
    try:
        r = _lib.make_request(0, 1, np.array([1.0, 1.0, 1.0]) * ARCMIN, hodset=a, maxeval=MAXEVAL)
        with pytest.raises(_lib.G3LError):
            srv.submit([r], tag=1)
        srv.submit([], tag=2)       # an empty request completes at once
        srv.step()
        assert srv.poll() == [(2, [])]
    finally:
        srv.close()
