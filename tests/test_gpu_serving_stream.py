"""The serving workload against the oracle: a seeded request stream
(g3lhalo_amd.workloads.serve_stream: prior HOD sets, two-population requests
at the three type pairs of calculateNNMap_twoPop.x, 1-200 triplets, random
NNMap/NMapMap term masks, maxEval 5e6) submitted to one server over a bounded
region pool while earlier requests are still refining, so integrals of
different requests share rounds and wait for pool memory.  A sample of every
request's integrals (tests/golden/serve_stream_oracle.npz) must match the
oracle -- the 1- and 2-halo terms and NMapMap to rounding with the oracle's
integrand-call count, the NNMap 3-halo term inside its error estimate -- and
two requests rerun alone through g3l_eval_batch must give the served results
bit for bit (a result does not depend on what else is in flight)."""
import os

import numpy as np
import pytest

from conftest import ROOT, T, say

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden", "serve_stream_oracle.npz")


def test_served_stream_matches_oracle(tables):
    from g3lhalo_amd import _lib
    from g3lhalo_amd import workloads as W
    g = np.load(GOLD)
    assert str(g["tables_checksum"]) == T.checksum(tables)
    stream = W.serve_stream(int(g["n_requests"]), seed=int(g["stream_seed"]))
    c = _lib.Context(tables)
    try:
        sids = [c.hodset(q["hod1"], q["hod2"], identical=not q["two"]) for q in stream]
        reqs = [[_lib.make_request(st, t, th * T.ARCMIN_IN_RAD, hodset=sid, type1=t1, type2=t2,
                                   reltol=1e-4, maxeval=5000000)
                 for (st, t, th, t1, t2) in W.request_integrals(q)]
                for q, sid in zip(stream, sids)]
        # a pool of 2^22 slots holds about 160 of these integrals at once
        srv = _lib.Server(c, pool_slots=1 << 22, max_integrals=1 << 14)
        done = {}
        waited = 0
        try:
            for i, r in enumerate(reqs):
                say(f"serving stream: request {i} ({len(r)} integrals) joins")
                srv.submit(r, tag=i)
                for _ in range(3):     # the next request arrives while these refine
                    st = srv.step()
                    waited = max(waited, st.pending)
                    done.update({tag: res for tag, res in srv.poll()})
            while len(done) < len(reqs):
                srv.step()
                done.update({tag: res for tag, res in srv.poll()})
        finally:
            srv.close()
        assert waited > 0   # the pool bound made integrals wait for admission
        say("serving stream: two requests alone through g3l_eval_batch")
        alone = {i: c.eval_batch(reqs[i]) for i in (0, len(reqs) - 1)}
    finally:
        c.close()
    for i, res in alone.items():
        for a, b in zip(res, done[i]):
            assert (a.value, a.err, a.neval, a.rounds) == (b.value, b.err, b.neval, b.rounds)
    for k in range(len(g["value"])):
        i, pos, stat, term = (int(g[f][k]) for f in ("request", "position", "stat", "term"))
        r = done[i][pos]
        d = abs(r.value - g["value"][k]) / abs(g["value"][k])
        print(f"request {i} integral {pos} (stat {stat} term {term}): rel {d:.2e}, rounds "
              f"{r.rounds} / {g['ncalls'][k]}")
        assert r.neval == g["neval"][k]
        if (stat, term) == (0, 3):
            assert abs(r.raw - g["raw"][k]) < 0.1 * g["err"][k]
        else:
            assert d < 1e-10 and r.rounds == g["ncalls"][k]
