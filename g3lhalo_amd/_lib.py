"""ctypes binding of the C ABI in include/g3l.h (libg3l_hip.so).

The library is built in-tree by ``python -m g3lhalo_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libg3l_hip.so")

G3L_OK, G3L_EINVAL, G3L_ENOMEM, G3L_EHIP, G3L_ENOMEM_POOL, G3L_MAXEVAL_HIT, G3L_NONFINITE, \
    G3L_ENODEV, G3L_ETOOBIG = range(9)
STATUS_NAMES = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EHIP", 4: "ENOMEM_POOL",
                5: "MAXEVAL_HIT", 6: "NONFINITE", 7: "ENODEV", 8: "ETOOBIG"}
NOALLOC_SCRATCH_BYTES = 64 << 20
NNMAP, NMAPMAP = 0, 1
FP64, MIXED = 0, 1
LAUNCH_STREAM, LAUNCH_GRAPH = 0, 1

_dp = C.POINTER(C.c_double)


class Tables(C.Structure):
    _fields_ = [("nbins", C.c_int32), ("zmin", C.c_double), ("zmax", C.c_double),
                ("kmin", C.c_double), ("kmax", C.c_double), ("mmin", C.c_double),
                ("mmax", C.c_double), ("g", _dp), ("p_lens1", _dp), ("p_lens2", _dp),
                ("w", _dp), ("dwdz", _dp), ("hmf", _dp), ("P_lin", _dp), ("b_h", _dp),
                ("concentration", _dp), ("H0", C.c_double), ("Omega_m", C.c_double),
                ("Omega_Lambda", C.c_double)]


class Hod(C.Structure):
    _fields_ = [("f", C.c_double), ("alpha", C.c_double), ("mth", C.c_double),
                ("sigma", C.c_double), ("mprime", C.c_double), ("beta", C.c_double)]


class Request(C.Structure):
    _fields_ = [("stat", C.c_int32), ("term", C.c_int32), ("hodset", C.c_int32),
                ("type1", C.c_int32), ("type2", C.c_int32), ("pad", C.c_int32),
                ("theta", C.c_double * 3), ("reltol", C.c_double), ("maxeval", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("value", C.c_double), ("raw", C.c_double), ("err", C.c_double),
                ("neval", C.c_uint64), ("rounds", C.c_uint32), ("status", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("eval_ms", C.c_double), ("select_ms", C.c_double), ("wall_ms", C.c_double),
                ("eval_launches", C.c_uint64), ("evals", C.c_uint64), ("rounds", C.c_uint32),
                ("pad", C.c_uint32), ("kind_ms", C.c_double * 8),
                ("kind_launches", C.c_uint64 * 8), ("kind_evals", C.c_uint64 * 8)]


class TuneReport(C.Structure):
    _fields_ = [("variant", C.c_int32 * 8), ("ms", (C.c_double * 4) * 8),
                ("group", (C.c_int32 * 4) * 8)]


class ServerConfig(C.Structure):
    _fields_ = [("pool_slots", C.c_uint64), ("max_integrals", C.c_uint32),
                ("precision", C.c_int32)]


class ServerStatus(C.Structure):
    _fields_ = [("in_flight", C.c_uint32), ("pending", C.c_uint32), ("completed", C.c_uint32),
                ("admitted", C.c_uint32), ("rounds", C.c_uint64)]


class ServerRequestInfo(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("submit_round", C.c_uint64), ("admit_round", C.c_uint64),
                ("done_round", C.c_uint64)]


class Cosmology(C.Structure):
    _fields_ = [("Omega_m", C.c_double), ("Omega_Lambda", C.c_double), ("Omega_b", C.c_double),
                ("sigma8", C.c_double), ("H0", C.c_double), ("ns", C.c_double)]


class PrecomputeIn(C.Structure):
    _fields_ = [("cosmo", Cosmology), ("zmin", C.c_double), ("zmax", C.c_double),
                ("mmin", C.c_double), ("mmax", C.c_double), ("kmin", C.c_double),
                ("kmax", C.c_double), ("nbins", C.c_int32), ("nz_z", _dp), ("nz_n", _dp),
                ("nz_count", C.c_int32)]


class PrecomputeOut(C.Structure):
    _fields_ = [(n, _dp) for n in ("z", "m", "k", "g", "w", "dwdz", "hmf", "P_lin", "b_h",
                                   "concentration")]


# Every symbol include/g3l.h declares, with its ctypes signature.
SIGNATURES = {
    "g3l_version": (C.c_char_p, []),
    "g3l_ctx_create": (C.c_int, [C.POINTER(Tables), C.POINTER(C.c_void_p)]),
    "g3l_ctx_destroy": (C.c_int, [C.c_void_p]),
    "g3l_hodset_create": (C.c_int, [C.c_void_p, C.POINTER(Hod), C.POINTER(Hod), C.c_double,
                                    C.c_double, C.c_int32, C.POINTER(C.c_int32)]),
    "g3l_hodset_create_many": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Hod), C.POINTER(Hod),
                                         _dp, _dp, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]),
    "g3l_hodset_densities": (C.c_int, [C.c_void_p, C.c_int32, _dp, _dp, _dp]),
    "g3l_hodset_clear": (C.c_int, [C.c_void_p]),
    "g3l_eval_plain": (C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(Result)]),
    "g3l_eval_batch": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_size_t, C.c_int32,
                                 C.POINTER(Result)]),
    "g3l_integrand": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_int32, C.c_size_t, _dp,
                                _dp]),
    "g3l_eval_batch_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_size_t,
                                         C.c_int32, C.POINTER(Result)]),
    "g3l_last_batch_stats": (C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    "g3l_set_trace": (C.c_int, [C.c_void_p, C.c_uint32]),
    "g3l_last_trace": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32), C.c_uint32,
                                 C.POINTER(C.c_uint32)]),
    "g3l_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(TuneReport)]),
    "g3l_set_variant": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "g3l_set_mixed_policy": (C.c_int, [C.c_void_p, C.c_uint32]),
    "g3l_get_mixed_policy": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "g3l_set_launch_mode": (C.c_int, [C.c_void_p, C.c_int32]),
    "g3l_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "g3l_server_submit": (C.c_int, [C.c_void_p, C.POINTER(Request), C.c_size_t,
                                    C.POINTER(Result), C.c_uint64]),
    "g3l_server_step": (C.c_int, [C.c_void_p, C.POINTER(ServerStatus)]),
    "g3l_server_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_size_t,
                                  C.POINTER(C.c_size_t)]),
    "g3l_server_poll_info": (C.c_int, [C.c_void_p, C.POINTER(ServerRequestInfo), C.c_size_t,
                                       C.POINTER(C.c_size_t)]),
    "g3l_server_destroy": (C.c_int, [C.c_void_p]),
    "g3l_precompute": (C.c_int, [C.POINTER(PrecomputeIn), C.POINTER(PrecomputeOut)]),
}

_lib = None


class G3LError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} failed: {STATUS_NAMES.get(rc, rc)}")
        self.rc = rc


def load():
    """Loads libg3l_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m g3lhalo_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn, rc):
    if rc != G3L_OK:
        raise G3LError(fn, rc)


def _p(a):
    return a.ctypes.data_as(_dp)


def hod_struct(h) -> Hod:
    h = [float(v) for v in h]
    return Hod(*h)


def make_request(stat, term, theta, hodset=0, type1=1, type2=2, reltol=1e-3, maxeval=5000000):
    r = Request()
    r.stat, r.term, r.hodset, r.type1, r.type2 = int(stat), int(term), int(hodset), int(type1), \
        int(type2)
    r.theta[0], r.theta[1], r.theta[2] = (float(t) for t in theta)
    r.reltol = float(reltol)
    r.maxeval = int(maxeval)
    return r


class Context:
    """One g3l_ctx: tables in HBM plus HOD sets."""

    def __init__(self, tables):
        L = load()
        self._arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in tables.arrays()]
        t = Tables()
        t.nbins = int(tables.nbins)
        t.zmin, t.zmax, t.kmin, t.kmax, t.mmin, t.mmax = (float(tables.zmin), float(tables.zmax),
                                                           float(tables.kmin), float(tables.kmax),
                                                           float(tables.mmin), float(tables.mmax))
        (t.g, t.p_lens1, t.p_lens2, t.w, t.dwdz, t.hmf, t.P_lin, t.b_h,
         t.concentration) = [_p(a) for a in self._arrs]
        t.H0 = float(tables.cosmo.H0)
        t.Omega_m = float(tables.cosmo.Omega_m)
        t.Omega_Lambda = float(tables.cosmo.Omega_Lambda)
        self.nbins = int(tables.nbins)
        h = C.c_void_p()
        check("g3l_ctx_create", L.g3l_ctx_create(C.byref(t), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            load().g3l_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hodset(self, hod1, hod2=None, A=0.0, eps=0.0, identical=False) -> int:
        a = hod_struct(hod1)
        b = hod_struct(hod2 if hod2 is not None else hod1)
        sid = C.c_int32()
        check("g3l_hodset_create", load().g3l_hodset_create(self.h, C.byref(a), C.byref(b),
                                                             float(A), float(eps),
                                                             int(identical), C.byref(sid)))
        return sid.value

    def hodsets(self, hod1s, hod2s=None, A=None, eps=None) -> int:
        n = len(hod1s)
        H1 = (Hod * n)(*[hod_struct(h) for h in hod1s])
        H2 = (Hod * n)(*[hod_struct(h) for h in (hod2s if hod2s is not None else hod1s)])
        Aa = np.ascontiguousarray(A if A is not None else np.zeros(n), dtype=np.float64)
        Ea = np.ascontiguousarray(eps if eps is not None else np.zeros(n), dtype=np.float64)
        first = C.c_int32()
        check("g3l_hodset_create_many",
              load().g3l_hodset_create_many(self.h, n, H1, H2, _p(Aa), _p(Ea), None,
                                            C.byref(first)))
        return first.value

    def densities(self, sid):
        n = self.nbins
        r, a, b = np.zeros(n), np.zeros(n), np.zeros(n)
        check("g3l_hodset_densities", load().g3l_hodset_densities(self.h, sid, _p(r), _p(a),
                                                                   _p(b)))
        return r, a, b

    def eval_plain(self, req: Request) -> Result:
        out = Result()
        check("g3l_eval_plain", load().g3l_eval_plain(self.h, C.byref(req), C.byref(out)))
        return out

    def eval_batch(self, reqs, precision=FP64):
        n = len(reqs)
        arr = (Request * n)(*reqs)
        out = (Result * n)()
        check("g3l_eval_batch", load().g3l_eval_batch(self.h, arr, n, int(precision), out))
        return list(out)

    def eval_batch_noalloc(self, reqs, precision=FP64):
        """g3l_eval_batch_noalloc: the batched engine over the fixed device
        scratch array (no allocation in the call)."""
        n = len(reqs)
        arr = (Request * max(n, 1))(*reqs)
        out = (Result * max(n, 1))()
        check("g3l_eval_batch_noalloc",
              load().g3l_eval_batch_noalloc(self.h, arr, n, int(precision), out))
        return list(out)[:n]

    def set_trace(self, cap: int):
        """Record the batch-size trace of following eval_batch calls (cap
        rounds per integral; 0 stops)."""
        check("g3l_set_trace", load().g3l_set_trace(self.h, int(cap)))

    def last_trace(self, i: int, cap: int = 4097) -> np.ndarray:
        """Regions per integrand call of integral i of the last traced
        eval_batch (the root region first)."""
        buf = (C.c_uint32 * cap)()
        n = C.c_uint32()
        check("g3l_last_trace", load().g3l_last_trace(self.h, int(i), buf, int(cap), C.byref(n)))
        return np.array(buf[:n.value], dtype=np.int64)

    def integrand(self, req: Request, x, precision=FP64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = np.zeros(x.shape[0])
        check("g3l_integrand", load().g3l_integrand(self.h, C.byref(req), int(precision),
                                                     x.shape[0], _p(x), _p(f)))
        return f

    def autotune(self, hodset=0, precision=FP64) -> dict:
        """Times the launch variants and keeps the fastest per kind; returns
        {kind: {"variant", "ms": [per variant], "group": [regions per block]}}."""
        r = TuneReport()
        check("g3l_autotune", load().g3l_autotune(self.h, int(hodset), int(precision),
                                                   C.byref(r)))
        nv = sum(1 for v in range(4) if r.group[1][v] > 0)
        return {k: {"variant": int(r.variant[k]), "ms": [round(r.ms[k][v], 3) for v in range(nv)],
                    "group": [int(r.group[k][v]) for v in range(nv)]} for k in range(1, 7)}

    def set_variant(self, kind, precision, variant):
        check("g3l_set_variant", load().g3l_set_variant(self.h, int(kind), int(precision),
                                                         int(variant)))

    def set_mixed_policy(self, fp32_kinds: int):
        """Kinds (bit k for kind k) that G3L_MIXED evaluates in FP32 bulk."""
        check("g3l_set_mixed_policy", load().g3l_set_mixed_policy(self.h, int(fp32_kinds)))

    def mixed_policy(self) -> int:
        v = C.c_uint32()
        check("g3l_get_mixed_policy", load().g3l_get_mixed_policy(self.h, C.byref(v)))
        return v.value

    def set_launch_mode(self, mode):
        """LAUNCH_GRAPH (default): one HIP graph launch per refinement round;
        LAUNCH_STREAM: the same kernels launched one by one."""
        check("g3l_set_launch_mode", load().g3l_set_launch_mode(self.h, int(mode)))

    def last_stats(self) -> BatchStats:
        s = BatchStats()
        check("g3l_last_batch_stats", load().g3l_last_batch_stats(self.h, C.byref(s)))
        return s


def precompute(cosmo: dict, zmin, zmax, mmin, mmax, kmin, kmax, nbins, nz_z, nz_n) -> dict:
    """g3l_precompute: the model's tables (z-major 2-D arrays) from a cosmology
    {Omega_m, Omega_Lambda, Omega_b, sigma8, H0, ns} and source n(z) samples."""
    z = np.ascontiguousarray(nz_z, dtype=np.float64)
    n = np.ascontiguousarray(nz_n, dtype=np.float64)
    cin = PrecomputeIn(Cosmology(*(float(cosmo[k]) for k in ("Omega_m", "Omega_Lambda", "Omega_b",
                                                              "sigma8", "H0", "ns"))),
                       zmin, zmax, mmin, mmax, kmin, kmax, int(nbins), _p(z), _p(n), len(z))
    out = {k: np.zeros(nbins) for k in ("z", "m", "k", "g", "w", "dwdz")}
    out.update({k: np.zeros((nbins, nbins)) for k in ("hmf", "P_lin", "b_h", "concentration")})
    cout = PrecomputeOut(*(_p(out[k]) for k in ("z", "m", "k", "g", "w", "dwdz", "hmf", "P_lin",
                                                 "b_h", "concentration")))
    check("g3l_precompute", load().g3l_precompute(C.byref(cin), C.byref(cout)))
    return out


class Server:
    """A g3l_server on a Context: submit requests, step, poll."""

    def __init__(self, ctx: "Context", pool_slots=0, max_integrals=0, precision=FP64):
        self.ctx = ctx
        self.h = C.c_void_p()
        cfg = ServerConfig(int(pool_slots), int(max_integrals), int(precision))
        check("g3l_server_create", load().g3l_server_create(ctx.h, C.byref(cfg),
                                                             C.byref(self.h)))
        self._out = {}   # tag -> (requests, results) kept alive until polled

    def submit(self, reqs, tag: int):
        n = len(reqs)
        arr = (Request * max(n, 1))(*reqs)
        out = (Result * max(n, 1))()
        self._out[int(tag)] = (arr, out, n)
        rc = load().g3l_server_submit(self.h, arr, n, out, int(tag))
        if rc != G3L_OK:
            self._out.pop(int(tag), None)
        check("g3l_server_submit", rc)

    def step(self) -> ServerStatus:
        st = ServerStatus()
        check("g3l_server_step", load().g3l_server_step(self.h, C.byref(st)))
        return st

    def poll(self, cap=1024):
        """Completed requests: list of (tag, [Result])."""
        return [(tag, res) for tag, res, _ in self.poll_info(cap)]

    def poll_info(self, cap=1024):
        """Completed requests: list of (tag, [Result], (submit_round,
        admit_round, done_round))."""
        info = (ServerRequestInfo * cap)()
        n = C.c_size_t()
        check("g3l_server_poll_info", load().g3l_server_poll_info(self.h, info, cap, C.byref(n)))
        done = []
        for i in range(n.value):
            tag = int(info[i].tag)
            arr, out, k = self._out.pop(tag)
            done.append((tag, [out[j] for j in range(k)],
                         (int(info[i].submit_round), int(info[i].admit_round),
                          int(info[i].done_round))))
        return done

    def close(self):
        if self.h:
            load().g3l_server_destroy(self.h)
            self.h = C.c_void_p()
