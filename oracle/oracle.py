"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path
(g3lhalo_amd) never imports this module.

``reference_hcubature()`` returns the reference's own ``hcubature_v``
compiled from ``src/hcubature.cpp`` into ``oracle/_ref`` (None when that
build is absent); passing it as ``hcub=`` runs the reference integrator over
the oracle's integrands.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_hcubature.so")

NNMAP, NMAPMAP = 0, 1

_dp = C.POINTER(C.c_double)


class Stats(C.Structure):
    _fields_ = [("neval", C.c_uint64), ("ncalls", C.c_uint32), ("trace_len", C.c_uint32)]


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "all"], check=True, stdout=subprocess.DEVNULL)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_model_create.restype = C.c_void_p
        L.orc_model_create.argtypes = [C.c_int] + [C.c_double] * 6 + [_dp] * 9 + [C.c_double] * 3
        L.orc_model_destroy.argtypes = [C.c_void_p]
        L.orc_model_set_hods.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int,
                                         C.c_void_p]
        L.orc_model_densities.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_model_borders.argtypes = [C.c_void_p, _dp]
        L.orc_eval_term.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_int, C.c_double, C.c_uint64,
                                    C.c_void_p, _dp, C.POINTER(Stats), C.POINTER(C.c_uint32),
                                    C.c_uint32]
        L.orc_integrand.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_int, C.c_int, C.c_size_t, _dp, _dp]
        L.orc_genz.argtypes = [C.c_int, C.c_uint, _dp, _dp, _dp, _dp, C.c_uint64, C.c_double,
                               C.c_void_p, _dp, _dp, C.POINTER(Stats)]
        L.orc_sici.argtypes = [C.c_double, _dp, _dp]
        for name in ("orc_aperture",):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        for name in ("orc_ncen", "orc_nsat"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [_dp, C.c_double]
        L.orc_nsatnsat.restype = C.c_double
        L.orc_nsatnsat.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_unfw.restype = C.c_double
        L.orc_unfw.argtypes = [C.c_void_p] + [C.c_double] * 4
        L.orc_bilin.restype = C.c_double
        L.orc_bilin.argtypes = [C.c_void_p] + [C.c_double] * 4
        L.orc_fkernel.restype = C.c_double
        L.orc_fkernel.argtypes = [C.c_double] * 3
        L.orc_points_per_region.restype = C.c_uint32
        L.orc_points_per_region.argtypes = [C.c_uint]
        L.orc_set_ulp_seed.restype = None
        L.orc_set_ulp_seed.argtypes = [C.c_uint64]
        _lib = L
    return _lib


def reference_hcubature():
    """Address of the reference's hcubature_v, or None if not built."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        _ref = C.CDLL(REF_PATH)
    return C.cast(_ref.hcubature_v, C.c_void_p).value


def set_ulp_seed(seed: int) -> None:
    """Sensitivity probe: a nonzero seed moves the integrands' exp, log, pow,
    sin, cos and erf results by at most one unit in the last place (a stand-in
    for another math library), except the nearest-index chains (masses 10^x
    and the table-row logarithms, whose region-centre arguments sit on
    rounding ties); 0 restores the exact functions."""
    lib().orc_set_ulp_seed(seed)


def _p(a):
    return a.ctypes.data_as(_dp)


def sici(x):
    si, ci = C.c_double(), C.c_double()
    lib().orc_sici(float(x), C.byref(si), C.byref(ci))
    return si.value, ci.value


def aperture(eta):
    return lib().orc_aperture(float(eta))


def ncen(hod, m):
    h = np.ascontiguousarray(hod, dtype=np.float64)
    return lib().orc_ncen(_p(h), float(m))


def nsat(hod, m):
    h = np.ascontiguousarray(hod, dtype=np.float64)
    return lib().orc_nsat(_p(h), float(m))


def nsatnsat(hod1, hod2, same, A, eps, m):
    a = np.ascontiguousarray(hod1, dtype=np.float64)
    b = np.ascontiguousarray(hod2, dtype=np.float64)
    return lib().orc_nsatnsat(_p(a), _p(b), int(same), float(A), float(eps), float(m))


def fkernel(k1, k2, c):
    return lib().orc_fkernel(float(k1), float(k2), float(c))


def genz(family, a, u, lo, hi, maxeval=0, reltol=1e-6, hcub=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    val, err, st = C.c_double(), C.c_double(), Stats()
    rc = lib().orc_genz(int(family), len(a), _p(a), _p(u), _p(lo), _p(hi), int(maxeval),
                        float(reltol), hcub, C.byref(val), C.byref(err), C.byref(st))
    if rc:
        raise RuntimeError(f"orc_genz failed rc={rc}")
    return val.value, err.value, st.neval, st.ncalls


class Model:
    """Oracle model over a g3lhalo_amd.tables.Tables instance."""

    def __init__(self, t, hcub=None):
        self.t = t
        self.hcub = hcub
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in t.arrays()]
        self._keep = arrs
        c = t.cosmo
        self.h = lib().orc_model_create(int(t.nbins), t.zmin, t.zmax, t.kmin, t.kmax, t.mmin,
                                        t.mmax, *[_p(a) for a in arrs], c.H0, c.Omega_m,
                                        c.Omega_Lambda)
        if not self.h:
            raise RuntimeError("orc_model_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_model_destroy(self.h)
            self.h = None

    def set_hods(self, hod1, hod2=None, A=0.0, eps=0.0, identical=False):
        a = np.ascontiguousarray(hod1, dtype=np.float64)
        b = np.ascontiguousarray(hod2 if hod2 is not None else hod1, dtype=np.float64)
        rc = lib().orc_model_set_hods(self.h, _p(a), _p(b), float(A), float(eps),
                                      int(identical), self.hcub)
        if rc:
            raise RuntimeError(f"orc_model_set_hods rc={rc}")

    def densities(self):
        n = self.t.nbins
        r, a, b = np.zeros(n), np.zeros(n), np.zeros(n)
        if lib().orc_model_densities(self.h, _p(r), _p(a), _p(b)):
            raise RuntimeError("densities not computed")
        return r, a, b

    def borders(self):
        o = np.zeros(4)
        lib().orc_model_borders(self.h, _p(o))
        return o

    def eval_term(self, stat, term, theta, type1=1, type2=2, reltol=1e-3, maxeval=5000000,
                  trace_cap=0, hcub="model"):
        """Returns dict(value, raw, err, neval, ncalls, trace)."""
        out = np.zeros(3)
        st = Stats()
        tr = np.zeros(max(trace_cap, 1), dtype=np.uint32)
        h = self.hcub if hcub == "model" else hcub
        rc = lib().orc_eval_term(self.h, int(stat), int(term), float(theta[0]), float(theta[1]),
                                 float(theta[2]), int(type1), int(type2), float(reltol),
                                 int(maxeval), h, _p(out), C.byref(st),
                                 tr.ctypes.data_as(C.POINTER(C.c_uint32)), int(trace_cap))
        if rc:
            raise RuntimeError(f"orc_eval_term rc={rc}")
        return {"value": out[0], "raw": out[1], "err": out[2], "neval": int(st.neval),
                "ncalls": int(st.ncalls), "trace": tr[:st.trace_len].copy()}

    def integrand(self, stat, term, theta, x, type1=1, type2=2):
        x = np.ascontiguousarray(x, dtype=np.float64)
        npts = x.shape[0]
        f = np.zeros(npts)
        rc = lib().orc_integrand(self.h, int(stat), int(term), float(theta[0]), float(theta[1]),
                                 float(theta[2]), int(type1), int(type2), npts, _p(x), _p(f))
        if rc:
            raise RuntimeError(f"orc_integrand rc={rc}")
        return f

    def unfw(self, k, m, z, f):
        return lib().orc_unfw(self.h, float(k), float(m), float(z), float(f))

    def bilin(self, k1, k2, cosphi, z):
        return lib().orc_bilin(self.h, float(k1), float(k2), float(cosphi), float(z))
