"""Graph-launched refinement rounds (G3L_LAUNCH_GRAPH, the default) against
the same kernels launched one by one on the stream (G3L_LAUNCH_STREAM): every
result bit for bit, over several graph runs with eager stream work and
synchronisations between them, for the batched engine (all six term kinds,
two HOD sets, and the n_bar integrals of new HOD sets, which run on the same
engine) and for the continuous-batching server."""
import numpy as np
import pytest

from conftest import ARCMIN, T, say

pytestmark = pytest.mark.gpu

MAXEVAL = 200000


def _requests(a, b):
    from g3lhalo_amd import _lib
    th1 = np.array([2.0, 3.0, 5.0]) * ARCMIN
    th2 = np.array([6.0, 6.0, 1.5]) * ARCMIN
    reqs = [_lib.make_request(st, t, th, hodset=a, reltol=1e-4, maxeval=MAXEVAL)
            for st in (0, 1) for th in (th1, th2) for t in (1, 2, 3)]
    reqs += [_lib.make_request(0, t, th2, hodset=b, type1=1, type2=2, reltol=1e-4,
                               maxeval=MAXEVAL) for t in (1, 2, 3)]
    reqs += [_lib.make_request(0, 1, th1, hodset=a, reltol=0.3, maxeval=0)]  # converges
    return reqs


def _key(res):
    return [(r.status, r.neval, r.value, r.err, r.rounds) for r in res]


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_rounds_equal_stream_rounds(tables, precision):
    from g3lhalo_amd import _lib
    say(f"graph vs stream rounds, precision {precision}")
    c = _lib.Context(tables)
    try:
        c.set_mixed_policy(0x7e)   # every term in the FP32-bulk form under MIXED
        hods = T.hod_prior_draws(4, seed=31)
        c.set_launch_mode(_lib.LAUNCH_STREAM)
        first_s = c.hodsets(hods, hods)
        a, b = first_s, first_s + 1
        reqs = _requests(a, b)
        ref = _key(c.eval_batch(reqs, precision=precision))
        st_stream = c.last_stats()
        c.set_launch_mode(_lib.LAUNCH_GRAPH)
        first_g = c.hodsets(hods, hods)   # n_bar on the graph path
        for s_id, g_id in ((first_s, first_g), (first_s + 3, first_g + 3)):
            ds, dg = c.densities(s_id), c.densities(g_id)
            for x, y in zip(ds, dg):
                assert np.array_equal(x, y)
        for rep in range(3):
            say(f"graph run {rep}")
            got = _key(c.eval_batch(reqs, precision=precision))
            assert got == ref
            st = c.last_stats()
            assert st.rounds == st_stream.rounds
            assert list(st.kind_evals) == list(st_stream.kind_evals)
            assert all(st.kind_ms[k] > 0 for k in range(1, 7))
            # eager work on the stream between graph runs
            c.set_launch_mode(_lib.LAUNCH_STREAM)
            assert _key(c.eval_batch(reqs[:3], precision=precision)) == ref[:3]
            c.set_launch_mode(_lib.LAUNCH_GRAPH)
    finally:
        c.close()


def test_server_graph_equals_stream(tables):
    from g3lhalo_amd import _lib
    say("server: graph vs stream rounds")
    c = _lib.Context(tables)
    try:
        a = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        b = c.hodset(T.FIDUCIAL_HOD, tuple(T.hod_prior_draws(1, seed=12)[0]), A=0.2, eps=0.1)
        reqs = _requests(a, b)
        out = {}
        for mode in (_lib.LAUNCH_STREAM, _lib.LAUNCH_GRAPH, _lib.LAUNCH_GRAPH):
            c.set_launch_mode(mode)
            srv = _lib.Server(c, pool_slots=1 << 21, max_integrals=1024)
            try:
                srv.submit(reqs[:7], tag=1)
                srv.step()
                srv.submit(reqs[7:], tag=2)
                res = {}
                for _ in range(5000):
                    srv.step()
                    for tag, r in srv.poll():
                        res[tag] = _key(r)
                    if len(res) == 2:
                        break
            finally:
                srv.close()
            assert len(res) == 2
            if 0 in out:
                assert res == out[0]
            out.setdefault(0, res)
        # the server's integrals alone in the batched engine
        c.set_launch_mode(_lib.LAUNCH_GRAPH)
        assert _key(c.eval_batch(reqs)) == out[0][1] + out[0][2]
    finally:
        c.close()
