"""Host-side checks of the C ABI (no device compute): the library loads,
exports every function include/g3l.h declares, and rejects bad arguments
with status codes instead of exiting."""
import ctypes as C
import os
import re

import pytest

from g3lhalo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "g3l.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(g3l_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 11
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with g3l.h"


def test_version():
    assert _lib.load().g3l_version().startswith(b"g3lhalo_amd")


def test_invalid_arguments_return_status():
    L = _lib.load()
    h = C.c_void_p()
    assert L.g3l_ctx_create(None, C.byref(h)) == _lib.G3L_EINVAL
    t = _lib.Tables()
    t.nbins = 1
    assert L.g3l_ctx_create(C.byref(t), C.byref(h)) == _lib.G3L_EINVAL
    t.nbins = 46341   # nbins^2 would overflow the tables' int indices
    assert L.g3l_ctx_create(C.byref(t), C.byref(h)) == _lib.G3L_EINVAL
    assert L.g3l_ctx_destroy(None) == _lib.G3L_EINVAL
    sid = C.c_int32()
    assert L.g3l_hodset_create(None, None, None, 0.0, 0.0, 0, C.byref(sid)) == _lib.G3L_EINVAL
    r = _lib.make_request(0, 1, (1e-4, 1e-4, 1e-4))
    out = _lib.Result()
    assert L.g3l_eval_plain(None, C.byref(r), C.byref(out)) == _lib.G3L_EINVAL
    assert L.g3l_eval_batch(None, C.byref(r), 1, 0, C.byref(out)) == _lib.G3L_EINVAL
    assert L.g3l_integrand(None, C.byref(r), 0, 0, None, None) == _lib.G3L_EINVAL
    st = _lib.BatchStats()
    assert L.g3l_last_batch_stats(None, C.byref(st)) == _lib.G3L_EINVAL


def test_struct_layouts():
    # the ctypes mirrors match the C layout (checked field by field offsets)
    assert C.sizeof(_lib.Request) == 4 * 6 + 8 * 3 + 8 + 8
    assert _lib.Request.theta.offset == 24
    assert C.sizeof(_lib.Result) == 8 * 4 + 4 + 4
    assert C.sizeof(_lib.Hod) == 48


def test_precompute_rejects_bad_arguments():
    """g3l_precompute validates its inputs before touching a device."""
    import numpy as np
    L = _lib.load()
    n = 4
    z = np.linspace(0.01, 3.0, 10)
    nz = np.ones(10)
    bufs = {k: np.zeros(n) for k in ("z", "m", "k", "g", "w", "dwdz")}
    bufs.update({k: np.zeros(n * n) for k in ("hmf", "P_lin", "b_h", "concentration")})
    out = _lib.PrecomputeOut(*(_lib._p(bufs[k]) for k in ("z", "m", "k", "g", "w", "dwdz", "hmf",
                                                           "P_lin", "b_h", "concentration")))

    def make(**kw):
        cfg = dict(zmin=0.001, zmax=2.0, mmin=1e10, mmax=1e17, kmin=1e-4, kmax=1e4, nbins=n,
                   Om=0.3, OL=0.7, Ob=0.045)
        cfg.update(kw)
        cos = _lib.Cosmology(cfg["Om"], cfg["OL"], cfg["Ob"], 0.8, 70.0, 0.96)
        return _lib.PrecomputeIn(cos, cfg["zmin"], cfg["zmax"], cfg["mmin"], cfg["mmax"],
                                 cfg["kmin"], cfg["kmax"], cfg["nbins"], _lib._p(z), _lib._p(nz),
                                 len(z))

    assert L.g3l_precompute(None, C.byref(out)) == _lib.G3L_EINVAL
    assert L.g3l_precompute(C.byref(make()), None) == _lib.G3L_EINVAL
    for bad in (dict(nbins=0), dict(zmax=0.0), dict(mmin=0.0), dict(mmax=1e9),
                dict(kmin=-1.0), dict(kmax=1e-5), dict(Ob=0.3), dict(Om=0.0)):
        assert L.g3l_precompute(C.byref(make(**bad)), C.byref(out)) == _lib.G3L_EINVAL, bad
    partial = _lib.PrecomputeOut()
    assert L.g3l_precompute(C.byref(make()), C.byref(partial)) == _lib.G3L_EINVAL
