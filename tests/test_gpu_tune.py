"""Launch variants (regions per block) give identical results, and the
autotuner keeps one of them per kind."""
import numpy as np
import pytest

from conftest import ARCMIN, T, say

pytestmark = pytest.mark.gpu

TERMS = [(0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3)]
NVARIANTS = 4  # block shapes per kind (regions per block)


@pytest.mark.parametrize("precision", [0, 1])
def test_variants_give_identical_results(tables, precision):
    from g3lhalo_amd import _lib
    c = _lib.Context(tables)
    try:
        c.set_mixed_policy(0x7e)   # every term in the FP32-bulk form under MIXED
        sid = c.hodset(T.FIDUCIAL_HOD, tuple(T.hod_prior_draws(1, seed=2)[0]))
        reqs = [_lib.make_request(s, t, np.array(th) * ARCMIN, hodset=sid, reltol=1e-4,
                                  maxeval=400000)
                for (s, t) in TERMS for th in ((1.5, 4.0, 9.0), (7.0, 7.0, 2.0))]
        out = []
        for v in range(NVARIANTS):
            say(f"variant {v} precision {precision}")
            for kind in range(1, 7):
                c.set_variant(kind, precision, v)
            out.append(c.eval_batch(reqs, precision=precision))
        for other in out[1:]:
            for a, b in zip(out[0], other):
                assert a.value == b.value and a.err == b.err and a.neval == b.neval
    finally:
        c.close()


def test_autotune_report(tables):
    from g3lhalo_amd import _lib
    c = _lib.Context(tables)
    try:
        sid = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        say("autotune FP64")
        rep = c.autotune(sid, _lib.FP64)
        print(rep)
        for k, r in rep.items():
            assert r["variant"] in range(len(r["ms"]))
            assert all(m > 0 for m in r["ms"])
            assert r["ms"][r["variant"]] == min(r["ms"])
            assert len(r["ms"]) == NVARIANTS and all(g > 0 for g in r["group"])
    finally:
        c.close()
