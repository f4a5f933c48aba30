#!/usr/bin/env python3
"""Benchmark: NNMap adaptive integrals per second on one MI355X.

Workload (default ``nnmap550``): NNMap 1h + 2h + 3h for the 550 aperture
triplets (10 log-spaced apertures 0.5-32 arcmin, theta1 <= theta2, 10 theta3),
fiducial HOD, relative tolerance 1e-4 for every term, 5e6 evaluations per
integral at most -- 1650 adaptive integrals of 5-7 dimensions.  One step is
one pass of the batched device-resident engine (g3l_eval_batch, FP64) over
all 1650 integrals; the tables are resident in HBM before timing starts.

Beside the batched value the line reports the plain one-integral-at-a-time
HIP path (g3l_eval_plain) on a fixed sample, the FP64 VALU roofline of the
evaluation kernel and the CPU oracle (FP64 C++/OpenMP restatement of the
reference CPU build) on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from g3lhalo_amd import _lib, flops  # noqa: E402
from g3lhalo_amd import tables as T  # noqa: E402
from g3lhalo_amd import workloads as W  # noqa: E402

FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X vector FP64
KIND_NAMES = {0: "n_bar", 1: "NNMap 1h", 2: "NNMap 2h", 3: "NNMap 3h", 4: "NMapMap 1h",
              5: "NMapMap 2h", 6: "NMapMap 3h"}


# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs of this workload scaled by a copy kernel of known size
# (profiles/k_eval_nnmap550_counters.md): 123.6 MB read + 14.6 MB written per
# k_eval<NNMap 2h, FP64> launch of the nnmap550 step.
TRAFFIC_BYTES_PER_LAUNCH = {("nnmap550", 2): 138210534}

# Hardware FP64 flops per evaluation, (kind, launch variant) -> 2 FMA + ADD +
# MUL + TRANS of the SQ_INSTS_VALU_*_F64 counters x 64 lanes / evaluations,
# from the same profile: the FP64 instructions the kernel issues (a lane
# masked off in a branch still counts), against the credited algorithmic
# work of flops.FLOPS_PER_EVAL.
HW_FP64_FLOPS_PER_EVAL = {(1, 0): 777, (2, 3): 1114, (3, 2): 1079}

# Algorithmic HBM bytes per region an evaluation launch handles: the
# region's centre, half-widths and volume (2D + 1 doubles) and its list item
# read; its value, error and split byte written.
def region_bytes(dim: int) -> int:
    return (2 * dim + 1) * 8 + 8 + 2 * 8 + 1


KIND_DIM = {0: 1, 1: 5, 2: 6, 3: 7, 4: 5, 5: 6, 6: 7}
KIND_PTS = {1: 15, 5: 93, 6: 149, 7: 241}


# Launch variant per kind when the autotuner does not run (the library's
# built-in choice); the kernel's template argument is kind + 16 x variant.
# the library's built-in launch variant per kind (used when not autotuned)
DEFAULT_VARIANT = {0: 0, 1: 0, 2: 3, 3: 2, 4: 0, 5: 3, 6: 2}


def kernel_kv(tuned: dict, prec: str, kind: int) -> int:
    v = tuned[prec][kind]["variant"] if prec in tuned else DEFAULT_VARIANT[kind]
    return kind + 16 * v


def workload(name: str, hodset: int = 0):
    reqs = []
    if name == "nnmap550":
        for th in T.triplets_550():
            for term in (1, 2, 3):
                reqs.append(_lib.make_request(0, term, th * T.ARCMIN_IN_RAD, hodset=hodset,
                                              type1=1, type2=2, reltol=1e-4, maxeval=5000000))
        desc = {"workload": "NNMap 1h+2h+3h, 550 triplets 0.5-32 arcmin, tol 1e-4, maxEval 5e6",
                "integrals": len(reqs), "nbins": 128, "hod": "fiducial", "types": "(1,2)"}
    elif name == "nmapmap550":
        for th in T.triplets_550():
            for term in (1, 2, 3):
                reqs.append(_lib.make_request(1, term, th * T.ARCMIN_IN_RAD, hodset=hodset,
                                              type1=1, type2=2, reltol=1e-4, maxeval=5000000))
        desc = {"workload": "NMapMap 1h+2h+3h, 550 triplets 0.5-32 arcmin, tol 1e-4, maxEval 5e6",
                "integrals": len(reqs), "nbins": 128, "hod": "fiducial"}
    elif name == "nnmap4":
        for th in T.triplets_equal_4():
            reqs.append(_lib.make_request(0, 1, th * T.ARCMIN_IN_RAD, hodset=hodset, type1=1,
                                          type2=2, reltol=1e-3, maxeval=5000000))
        desc = {"workload": "NNMap 1h, 4 equal-aperture triplets 1-8 arcmin, tol 1e-3",
                "integrals": len(reqs), "nbins": 128}
    else:
        raise SystemExit(f"unknown workload {name}")
    return reqs, desc


def cpu_model() -> str:
    """The host CPU's model name (the CPU baselines run on it)."""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tables, reqs, seconds_budget: float):
    """Oracle (FP64 C++/OpenMP port of the reference CPU build) on a sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    m = O.Model(tables)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    # an even spread over the workload, cut off once the budget is spent
    idx = np.linspace(0, len(reqs) - 1, 96).astype(int)
    done, t_used = 0, 0.0
    vals = {}
    for i in idx:
        q = reqs[i]
        t0 = time.perf_counter()
        o = m.eval_term(q.stat, q.term, list(q.theta), q.type1, q.type2, q.reltol, q.maxeval)
        t_used += time.perf_counter() - t0
        vals[int(i)] = o["value"]
        done += 1
        if t_used > seconds_budget:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": done / t_used, "unit": "integrals/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{done} of 96 evenly spaced integrals of the workload (first index "
                      f"{min(vals)}, last {max(vals)}), full maxEval, until {seconds_budget:g} s",
            "seconds": round(t_used, 2)}, vals


def serve(ctx, tables, n: int, rate: float, precision=_lib.FP64, seed: int = 4):
    """Continuous batching of a Poisson request stream through g3l_server_*:
    each request's HOD set (device n_bar) is made when it arrives, its
    integrals join the ones in flight at the next refinement round.  Latency
    runs from the scheduled arrival to the end of the step that completes the
    request, split at the start of the step that admitted its last integral
    into admission wait (queueing for pool memory and the host loop) and
    refinement."""
    stream = W.serve_stream(n, seed)
    rng = np.random.default_rng(seed + 2)
    arrive = np.cumsum(rng.exponential(1.0 / rate, size=n))
    srv = _lib.Server(ctx, pool_slots=1 << 26, max_integrals=1 << 17, precision=precision)
    finish = np.full(n, np.nan)
    admitted = np.full(n, np.nan)
    round_start = {}
    rounds_done = 0
    nint = 0
    i = 0
    t0 = time.perf_counter()
    try:
        while np.isnan(finish).any():
            now = time.perf_counter() - t0
            while i < n and arrive[i] <= now:
                q = stream[i]
                sid = ctx.hodset(q["hod1"], q["hod2"], identical=not q["two"])
                reqs = [_lib.make_request(st, t, th * T.ARCMIN_IN_RAD, hodset=sid, type1=t1,
                                          type2=t2, reltol=1e-4, maxeval=5000000)
                        for (st, t, th, t1, t2) in W.request_integrals(q)]
                nint += len(reqs)
                srv.submit(reqs, tag=i)
                i += 1
            ts = time.perf_counter() - t0
            st = srv.step()
            t = time.perf_counter() - t0
            if st.rounds > rounds_done:
                round_start[rounds_done] = ts
                rounds_done = st.rounds
            for tag, _, (_, adm, _) in srv.poll_info():
                finish[tag] = t
                admitted[tag] = round_start.get(adm, ts)
            if st.in_flight == 0 and st.pending == 0 and i < n:
                time.sleep(max(0.0, arrive[i] - (time.perf_counter() - t0)))
    finally:
        srv.close()
    lat = (finish - arrive) * 1e3
    wait = (admitted - arrive) * 1e3
    refine = (finish - admitted) * 1e3
    span = float(np.nanmax(finish) - arrive[0])

    def pct(a):
        return {"p50": round(float(np.percentile(a, 50)), 1),
                "p99": round(float(np.percentile(a, 99)), 1)}
    return {"value": round(nint / span, 3), "unit": "integrals/s", "requests": n,
            "integrals": nint, "arrival_rate_per_s": rate,
            "p50_ms": pct(lat)["p50"], "p99_ms": pct(lat)["p99"],
            "admission_wait_ms": pct(wait), "refinement_ms": pct(refine),
            "span_s": round(span, 3), "rounds": rounds_done,
            "stream": "per request: prior HOD (20% two-population with the three type pairs), "
                      "1-200 of the 550 triplets, random NNMap/NMapMap 1h/2h/3h mask; tol 1e-4, "
                      "maxEval 5e6"}


def vs_fixture(values, name: str, tables):
    """Per-term relative error of 550-triplet results (1h, 2h, 3h interleaved)
    against the oracle's committed values for the same workload, or None if
    the fixture is absent or was made from other tables."""
    path = os.path.join(ROOT, "tests", "golden", f"{name}_oracle.npz")
    if not os.path.exists(path):
        return None
    g = np.load(path)
    if str(g["tables_checksum"]) != T.checksum(tables) or len(g["value"]) != len(values):
        return None
    d = np.abs(np.asarray(values) - g["value"]) / np.abs(g["value"])
    return {f"{t}h": {"p50": float(np.median(d[t - 1::3])), "max": float(d[t - 1::3].max()),
                      "frac_above_1e-4": float(np.mean(d[t - 1::3] > 1e-4))} for t in (1, 2, 3)}


def hod_sweep(ctx, tables, n_hods: int = 256, pool_slots: int = 1 << 26, precision=_lib.FP64):
    """The batched fit sweep: n_hods HOD vectors from the uniform prior x the
    55 triplets (a_i, a_j, a_j) x NNMap 1h/2h/3h, every integral at tol 1e-4
    and maxEval 5e6, submitted at once to one server whose region pool is
    bounded (pool_slots slots, 249 B each): integrals are admitted as pool
    segments free up and every admitted integral refines each round.  The HOD
    sets' n_bar(z) are computed on the device inside the timed region."""
    hods = T.hod_prior_draws(n_hods, seed=20220406)
    trips = T.triplets_55()
    t0 = time.perf_counter()
    first = ctx.hodsets(hods, hods)
    srv = _lib.Server(ctx, pool_slots=pool_slots, max_integrals=1 << 17, precision=precision)
    nint, done, evals = 0, 0, 0
    lat = []
    try:
        for h in range(n_hods):
            reqs = [_lib.make_request(0, term, th * T.ARCMIN_IN_RAD, hodset=first + h, type1=1,
                                      type2=2, reltol=1e-4, maxeval=5000000)
                    for th in trips for term in (1, 2, 3)]
            nint += len(reqs)
            srv.submit(reqs, tag=h)
        rounds = 0
        first = {}
        while done < n_hods:
            srv.step()
            rounds += 1
            for tag, res in srv.poll():
                done += 1
                lat.append(time.perf_counter() - t0)
                evals += sum(int(r.neval) for r in res)
                if tag < 8:
                    first[tag] = [r.value for r in res]
    finally:
        srv.close()
    span = time.perf_counter() - t0
    parity = None
    path = os.path.join(ROOT, "tests", "golden", "hod_sweep8_oracle.npz")
    if os.path.exists(path) and len(first) == 8:
        g = np.load(path)
        v = np.concatenate([first[h] for h in range(8)])
        if str(g["tables_checksum"]) == T.checksum(tables) and len(v) == len(g["value"]):
            d = np.abs(v - g["value"]) / np.abs(g["value"])
            parity = {"integrals": int(len(d)), "hod_sets": 8,
                      "max_rel": {f"{t}h": float(d[t - 1::3].max()) for t in (1, 2, 3)},
                      "frac_above_1e-4": {f"{t}h": float(np.mean(d[t - 1::3] > 1e-4))
                                          for t in (1, 2, 3)},
                      "note": "the first 8 HOD sets of the sweep (all 55 triplets x 3 terms) "
                              "against the oracle's values (tests/golden/hod_sweep8_oracle.npz)"}
    return {"value": round(nint / span, 3), "unit": "integrals/s", "hod_sets": n_hods,
            "parity_vs_oracle": parity,
            "integrals": nint, "seconds": round(span, 3), "rounds": rounds,
            "evals": evals, "pool_slots": pool_slots,
            "pool_gib": round(pool_slots * 249 / 2**30, 2),
            "first_hod_done_s": round(min(lat), 3), "last_hod_done_s": round(max(lat), 3),
            "workload": f"{n_hods} prior HOD vectors x 55 triplets (a_i, a_j, a_j) x NNMap "
                        "1h/2h/3h, tol 1e-4, maxEval 5e6; device n_bar(z) of every HOD set "
                        "inside the timed region; one server, bounded region pool"}


def nb256_line():
    """The nnmap550 workload on 256-bin tables (every table on a 256-point
    grid, the reference's one bin count for all tables), FP64 batched, one
    warm-up and one timed step; a sample checked against the oracle's values
    on the same tables (tests/golden/nnmap550_nb256_oracle.npz)."""
    t = T.make_tables(nbins=256)
    c = _lib.Context(t)
    try:
        sid = c.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
        reqs, desc = workload("nnmap550", sid)
        c.eval_batch(reqs)
        t0 = time.perf_counter()
        res = c.eval_batch(reqs)
        dt = time.perf_counter() - t0
    finally:
        c.close()
    out = {"value": round(len(reqs) / dt, 3), "unit": "integrals/s",
           "ms_per_step": round(dt * 1e3, 3), "steps": 1, "warmup": 1, **desc, "nbins": 256}
    path = os.path.join(ROOT, "tests", "golden", "nnmap550_nb256_oracle.npz")
    if os.path.exists(path):
        g = np.load(path)
        if str(g["tables_checksum"]) == T.checksum(t):
            idx = 3 * g["triplet"] + g["term"] - 1
            d = np.abs(np.array([res[i].value for i in idx]) - g["value"]) / np.abs(g["value"])
            out["parity_vs_oracle"] = {
                "integrals": int(len(d)), "max_rel_1h_2h": float(d[g["term"] < 3].max()),
                "max_rel_3h": float(d[g["term"] == 3].max()),
                "sample": "triplets 0, 183, 366, 549 x 1h/2h/3h"}
    return out


def precompute_line(nbins: int = 128, cpu_nbins: int = 32):
    """The table precomputation (g3l_precompute) on the nnmap550 grids: the
    fiducial cosmology, z in [0.001, 2], m in [1e10, 1e17], k in [1e-4, 1e4],
    a source n(z) ~ z^2 exp(-(z/0.5)^1.5) on 100 points; beside it the
    reference's own doPrecomputations.x (built from its sources into
    oracle/_ref) on the host cores at a smaller grid, when present."""
    import subprocess
    import tempfile
    c = T.Cosmo()
    cosmo = {"Omega_m": c.Omega_m, "Omega_Lambda": c.Omega_Lambda, "Omega_b": c.Omega_b,
             "sigma8": c.sigma8, "H0": c.H0, "ns": c.ns}
    z = np.linspace(0.01, 3.0, 100)
    n = z ** 2 * np.exp(-(z / 0.5) ** 1.5)
    n /= n.sum() * (z[-1] - z[0]) / len(z)
    grid = (0.001, 2.0, 1e10, 1e17, 1e-4, 1e4)

    def integrals(nb):
        return 1 + nb + len(z) + 2 * nb * nb + 256 * nb

    _lib.precompute(cosmo, *grid, 8, z, n)   # warm-up
    t0 = time.perf_counter()
    _lib.precompute(cosmo, *grid, nbins, z, n)
    dt = time.perf_counter() - t0
    line = {"nbins": nbins, "seconds": round(dt, 3), "adaptive_integrals": integrals(nbins),
            "integrals_per_s": round(integrals(nbins) / dt, 1),
            "note": "all seven tables; the integrals refine on the device's batched GK15 engine"}
    ref = os.path.join(ROOT, "oracle", "_ref", "doPrecomputations.x")
    if os.path.exists(ref):
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(d, "cosmo.param"), "w") as f:
                f.write("".join(f"{k} {v!r}\n" for k, v in cosmo.items()))
            with open(os.path.join(d, "nz.dat"), "w") as f:
                f.write("".join(f"{a:.17g} {b:.17g}\n" for a, b in zip(z, n)))
            args = [ref, "cosmo.param", "nz.dat", *[repr(v) for v in grid], str(cpu_nbins),
                    "g.dat", "w.dat", "dwdz.dat", "hmf.dat", "pk.dat", "bh.dat", "conc.dat"]
            t0 = time.perf_counter()
            subprocess.run(args, cwd=d, check=True, capture_output=True,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
            cdt = time.perf_counter() - t0
        line["cpu_reference"] = {"nbins": cpu_nbins, "seconds": round(cdt, 3),
                                 "adaptive_integrals": integrals(cpu_nbins),
                                 "integrals_per_s": round(integrals(cpu_nbins) / cdt, 1),
                                 "cores": threads, "kind": "reference", "cpu": cpu_model()}
    return line


def copy_probe(mib: int) -> int:
    """One elementwise float32 kernel that reads mib MiB and writes mib MiB: the
    known-size reference against which the HBM traffic counters are scaled."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.add(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4}))
    return 0


def progress(msg: str) -> None:
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="nnmap550")
    ap.add_argument("--plain-sample", type=int, default=30,
                    help="integrals of the workload timed through the plain form")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests in the served run (0 skips it)")
    ap.add_argument("--serve-rates", default="4,8,12",
                    help="mean Poisson arrival rates of the served runs below capacity, "
                         "requests/s (comma-separated; the first is the 'served' line)")
    ap.add_argument("--sweep-hods", type=int, default=256,
                    help="HOD vectors in the batched fit sweep (0 skips it)")
    ap.add_argument("--no-nmapmap", action="store_true",
                    help="skip the NMapMap 550-triplet line")
    ap.add_argument("--no-precompute", action="store_true",
                    help="skip the table-precomputation line")
    ap.add_argument("--no-nb256", action="store_true",
                    help="skip the 256-bin variant of the workload")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run an elementwise kernel that reads and writes this many MiB "
                         "(calibrates the FETCH_SIZE / WRITE_SIZE counters) and exit")
    a = ap.parse_args(argv)
    if a.copy_probe_mib:
        return copy_probe(a.copy_probe_mib)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the integrals are independent and shard by request "
              "with no exchange; run it with --gpus 1", file=sys.stderr)
        return 2

    tables = T.make_tables()
    ctx = _lib.Context(tables)
    sid = ctx.hodset(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    reqs, desc = workload(a.workload, sid)
    # block shapes per integrand kind, timed on this device (results do not
    # depend on them); --no-autotune keeps the built-in choices (profiling
    # runs, so that every launch of a kernel belongs to the timed steps)
    tuned = {}
    if not a.no_autotune:
        tuned["fp64"] = ctx.autotune(sid, _lib.FP64)
        if not a.fp64_only:
            tuned["mixed"] = ctx.autotune(sid, _lib.MIXED)

    def run_form(precision):
        for _ in range(a.warmup):
            ctx.eval_batch(reqs, precision=precision)
        step_ms = []
        kind_ms = np.zeros(8)
        kind_launches = np.zeros(8, dtype=np.int64)
        kind_evals = np.zeros(8, dtype=np.int64)
        res = None
        parts = []
        t_start = time.perf_counter()
        for _ in range(a.steps):
            t0 = time.perf_counter()
            res = ctx.eval_batch(reqs, precision=precision)
            step_ms.append((time.perf_counter() - t0) * 1e3)
            st = ctx.last_stats()
            parts.append((st.eval_ms, st.select_ms, st.wall_ms, st.rounds))
            kind_ms += np.array(st.kind_ms[:])
            kind_launches += np.array(st.kind_launches[:], dtype=np.int64)
            kind_evals += np.array(st.kind_evals[:], dtype=np.int64)
        elapsed = time.perf_counter() - t_start
        ev, sel, wall, rounds = (float(np.mean([p[i] for p in parts])) for i in range(4))
        # where one step's time goes: evaluation launches, selection kernels
        # (both from event nodes of the round graphs) and the rest (host
        # turnaround between rounds, graph launches, the result copy)
        breakdown = {"eval": round(ev, 3), "select": round(sel, 3),
                     "rest": round(wall - ev - sel, 3), "rounds": int(rounds)}
        return {"value": len(reqs) * a.steps / elapsed, "step_ms": step_ms, "results": res,
                "kind_ms": kind_ms, "kind_launches": kind_launches, "kind_evals": kind_evals,
                "breakdown": breakdown}

    progress("batched FP64 steps")
    f64 = run_form(_lib.FP64)
    progress("batched mixed steps")
    mixed = run_form(_lib.MIXED) if not a.fp64_only else None
    head = f64
    value = head["value"]
    step_ms = head["step_ms"]
    results = f64["results"]
    bad = [r.status for r in results if r.status not in (_lib.G3L_OK, _lib.G3L_MAXEVAL_HIT)]
    kind_ms, kind_launches, kind_evals = head["kind_ms"], head["kind_launches"], head["kind_evals"]

    progress("plain sample")
    # plain one-integral-at-a-time HIP path on a fixed sample
    pidx = np.linspace(0, len(reqs) - 1, max(1, a.plain_sample)).astype(int)
    t0 = time.perf_counter()
    plain = {int(i): ctx.eval_plain(reqs[i]) for i in pidx}
    plain_s = time.perf_counter() - t0
    plain_rate = len(pidx) / plain_s
    dev = max(abs(plain[i].value - results[i].value) / max(abs(plain[i].value), 1e-300)
              for i in pidx)

    # roofline of the dominant evaluation kernel (largest share of device time):
    # algorithmic flops = its evaluations x frozen flops per evaluation
    per_kind = {}
    for k in range(7):
        if kind_launches[k] == 0:
            continue
        fl = float(kind_evals[k]) * flops.FLOPS_PER_EVAL[k]
        per_kind[KIND_NAMES[k]] = {
            "ms_total": round(float(kind_ms[k]), 3), "launches": int(kind_launches[k]),
            "avg_launch_ms": round(float(kind_ms[k] / kind_launches[k]), 4),
            "evals": int(kind_evals[k]), "flops_per_eval": flops.FLOPS_PER_EVAL[k],
            "tflops": round(fl / (kind_ms[k] * 1e-3) / 1e12, 3)}
    dom = int(np.argmax(kind_ms[:7]))
    dom_flops = float(kind_evals[dom]) * flops.FLOPS_PER_EVAL[dom]
    achieved = dom_flops / (kind_ms[dom] * 1e-3) / 1e12 if kind_ms[dom] > 0 else 0.0
    peak = FP64_VALU_PEAK_TFLOPS
    dom_var = int(tuned["fp64"][dom]["variant"]) if "fp64" in tuned else DEFAULT_VARIANT[dom]
    hw_fpe = HW_FP64_FLOPS_PER_EVAL.get((dom, dom_var))
    hw_fp64 = None
    if hw_fpe is not None and kind_ms[dom] > 0:
        hw_t = float(kind_evals[dom]) * hw_fpe / (kind_ms[dom] * 1e-3) / 1e12
        hw_fp64 = {"flops_per_eval": hw_fpe, "tflops": round(hw_t, 3),
                   "frac": round(hw_t / peak, 4),
                   "source": "SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 counters of this kernel "
                             "(profiles/k_eval_nnmap550_counters.md)"}
    dim = KIND_DIM[dom]
    regions_per_launch = (float(kind_evals[dom]) / KIND_PTS[dim] / kind_launches[dom]
                          if kind_launches[dom] else 0.0)
    alg_bytes = int(round(regions_per_launch * region_bytes(dim)))
    mixed_info = None
    if mixed is not None:
        mdev = np.array([abs(m.value - r.value) / max(abs(r.value), 1e-300)
                         for m, r in zip(mixed["results"], f64["results"])])
        terms = np.array([r.term for r in reqs])
        mk = int(np.argmax(mixed["kind_ms"][:7]))
        mflops = float(mixed["kind_evals"][mk]) * flops.FLOPS_PER_EVAL[mk]
        mach = mflops / (mixed["kind_ms"][mk] * 1e-3) / 1e12 if mixed["kind_ms"][mk] > 0 else 0.0
        mixed_info = {
            "value": round(mixed["value"], 3), "unit": "integrals/s",
            "ms_per_step": round(float(np.mean(mixed["step_ms"])), 3),
            "rel_diff_vs_fp64_per_integral": {
                "p50": float(np.quantile(mdev, 0.5)), "p90": float(np.quantile(mdev, 0.9)),
                "p99": float(np.quantile(mdev, 0.99)), "max": float(mdev.max()),
                "frac_above_1e-4": float(np.mean(mdev > 1e-4))},
            "rel_diff_vs_fp64_per_term": {
                f"{t}h": {"p50": float(np.quantile(mdev[terms == t], 0.5)),
                          "max": float(mdev[terms == t].max()),
                          "frac_above_1e-4": float(np.mean(mdev[terms == t] > 1e-4))}
                for t in (1, 2, 3)},
            "roofline": {"kernel": f"k_eval<{kernel_kv(tuned, 'mixed', mk)},PMix> "
                                   f"({KIND_NAMES[mk]})",
                         "achieved": round(mach, 3), "peak": round(flops.mixed_peak_tflops(mk), 2),
                         "unit": "TFLOP/s",
                         "frac": round(mach / flops.mixed_peak_tflops(mk), 4),
                         "peak_note": "FP32 at the rate of the unpacked instructions emitted "
                                      "(78.6 TFLOP/s, as FP64); with packed FP32 the ceiling "
                                      f"would be {flops.mixed_peak_tflops(mk, True):.1f}"},
            "note": "FP32 transcendental bulk; not the headline: these integrals end at the "
                    "evaluation cap unconverged, and FP32 noise re-orders near-tied regions"}

    if a.workload == "nnmap550":
        fx = {"fp64": vs_fixture([r.value for r in f64["results"]], "nnmap550", tables)}
        if mixed is not None:
            fx["mixed"] = vs_fixture([r.value for r in mixed["results"]], "nnmap550", tables)
        if mixed_info is not None and fx.get("mixed") is not None:
            mixed_info["rel_err_vs_oracle_per_term"] = fx["mixed"]
    line = {
        "metric": "NNMap integrals/sec at 1e-4 rel tol (batched vs plain, 1 GPU); p50/p99 latency",
        "value": round(value, 3),
        "unit": "integrals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(float(np.mean(step_ms)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic tables (EH no-wiggle P(k), Sheth-Tormen, NFW c(m,z)) on a 128-bin grid",
        "config": dict(desc, form="batched device-resident, FP64"),
        "mixed": mixed_info,
        "plain": {"value": round(plain_rate, 3), "unit": "integrals/s",
                  "sample": f"{len(pidx)} integrals of the workload",
                  "speedup_batched_over_plain": round(value / plain_rate, 2),
                  "speedup_mixed_over_plain": (round(mixed["value"] / plain_rate, 2)
                                               if mixed is not None else None),
                  "max_rel_diff_batched_vs_plain": float(dev)},
        "step_breakdown_ms": head["breakdown"],
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 3),
                       "p99": round(float(np.percentile(step_ms, 99)), 3),
                       "note": "batched form: every integral of a step completes with the step"},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                     "compute": "FP64 vector ALU issue (no MFMA: no contraction); MI355X FP64 "
                                "vector peak 78.6 TFLOP/s",
                     "peak": round(peak, 2), "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4),
                     "hw_fp64": hw_fp64,
                     "traffic": TRAFFIC_BYTES_PER_LAUNCH.get((a.workload, dom)),
                     "algorithmic_bytes": alg_bytes,
                     "traffic_over_algorithmic": (
                         round(TRAFFIC_BYTES_PER_LAUNCH[(a.workload, dom)] / alg_bytes, 2)
                         if (a.workload, dom) in TRAFFIC_BYTES_PER_LAUNCH and alg_bytes else None),
                     "kernel": f"k_eval<{kernel_kv(tuned, 'fp64', dom)},P64> "
                               f"({KIND_NAMES[dom]}, fused points + integrand + rule)",
                     "avg_launch_ms": round(float(kind_ms[dom] / max(kind_launches[dom], 1)), 4),
                     "per_kernel": per_kind},
        "autotune": {p: {KIND_NAMES[k]: r["group"][r["variant"]] for k, r in t.items()}
                     for p, t in tuned.items()},
        "rel_err_vs_oracle_per_term": (fx["fp64"] if a.workload == "nnmap550" else None),
        "evals_per_step": int(sum(int(r.neval) for r in results)),
        "failed_integrals": len(bad),
    }
    # device n_bar(z) for a sweep of HOD vectors (the set-up of a fit step):
    # 2 x 128 GK15 integrals per HOD pair, all sets in one call
    hods = T.hod_prior_draws(256, seed=20220405)
    t0 = time.perf_counter()
    ctx.hodsets(hods, hods)
    nbar_ms = (time.perf_counter() - t0) * 1e3
    line["nbar_sweep"] = {"hod_sets": 256, "integrals": 256 * 2 * tables.nbins,
                          "ms": round(nbar_ms, 3),
                          "integrals_per_s": round(256 * 2 * tables.nbins / (nbar_ms * 1e-3), 1)}
    if not a.no_nmapmap and a.workload == "nnmap550":
        # the same grid for the galaxy-shear-shear statistic, FP64 batched
        nreqs, ndesc = workload("nmapmap550", sid)
        ctx.eval_batch(nreqs)
        t0 = time.perf_counter()
        nres = ctx.eval_batch(nreqs)
        dt = time.perf_counter() - t0
        line["nmapmap550"] = {"value": round(len(nreqs) / dt, 3), "unit": "integrals/s",
                              "ms_per_step": round(dt * 1e3, 3), "steps": 1, "warmup": 1,
                              "evals": int(sum(int(r.neval) for r in nres)), **ndesc,
                              "rel_err_vs_oracle_per_term":
                                  vs_fixture([r.value for r in nres], "nmapmap550", tables)}
        if not a.fp64_only:
            ctx.eval_batch(nreqs, precision=_lib.MIXED)
            t0 = time.perf_counter()
            mres = ctx.eval_batch(nreqs, precision=_lib.MIXED)
            mdt = time.perf_counter() - t0
            line["nmapmap550"]["mixed"] = {
                "value": round(len(nreqs) / mdt, 3), "unit": "integrals/s",
                "rel_err_vs_oracle_per_term": vs_fixture([r.value for r in mres], "nmapmap550",
                                                         tables)}
    progress("n_bar sweep, NMapMap-550 done")
    if not a.no_nb256 and a.workload == "nnmap550":
        progress("256-bin variant")
        line["nnmap550_nb256"] = nb256_line()
    if a.sweep_hods > 0:
        progress("HOD sweep")
        line["hod_sweep"] = hod_sweep(ctx, tables, a.sweep_hods)
    if not a.no_precompute:
        progress("table precomputation")
        line["precompute"] = precompute_line()
    if a.serve_requests > 0:
        # below capacity (a request averages ~400 integrals, the device
        # refines ~4600 integrals/s: ~11 requests/s), then at the serving
        # configuration's nominal 200 requests/s (overload: latency is
        # queueing, the throughput is the device's capacity)
        progress("served runs")
        rates = [float(r) for r in a.serve_rates.split(",") if r]
        line["served"] = serve(ctx, tables, a.serve_requests, rates[0])
        line["served_at_rates"] = [line["served"]] + [serve(ctx, tables, a.serve_requests, r)
                                                      for r in rates[1:]]
        line["latency_ms"]["served_p50"] = line["served"]["p50_ms"]
        line["latency_ms"]["served_p99"] = line["served"]["p99_ms"]
        nominal = serve(ctx, tables, 2 * a.serve_requests, 200.0, seed=7)
        nominal["note"] = ("Poisson arrivals at the nominal 200/s: the stream outruns the device "
                           "and waits for pool memory; throughput at saturation")
        line["served_nominal_rate"] = nominal
    if not a.no_cpu:
        progress("CPU baseline")
        cb, ovals = cpu_baseline(tables, reqs, a.cpu_seconds)
        line["cpu_baseline"] = cb
        # the same integrals from the FP64 batched step against the oracle
        d = np.array([abs(results[i].value - v) / max(abs(v), 1e-300) for i, v in ovals.items()])
        line["parity_vs_oracle"] = {
            "integrals": len(d), "max_rel": float(d.max()), "median_rel": float(np.median(d)),
            "frac_above_1e-4": float(np.mean(d > 1e-4)),
            "note": "the CPU baseline's sample, full maxEval: batched FP64 values vs the "
                    "oracle's for the same integrals"}
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
