"""Checks that the mathematics fixes for the six term integrands' compositions
(src/NNMap_Model.cu:544-593, 787-856, 1003-1037; src/NMapMap_Model.cu:263-301,
402-468, 593-632), independent of how any implementation restates them.
Shared by the oracle's CPU test (tests/test_oracle_compositions.py) and the
device's (tests/test_gpu_compositions.py); not collected by itself.

Each integrand is a product of table look-ups, HOD occupations, NFW
transforms and n_bar(z) normalisations.  Doubling one input table changes a
product of powers of it by an exact power of two (multiplication by 2 is exact
in binary floating point), and n_bar(z) -- an adaptive integral of
m hmf (Ncen + Nsat) -- doubles exactly with hmf, because every rule sum, error
estimate and convergence decision of its Gauss-Kronrod refinement scales
with it.  So the exponent of 2 each term picks up is fixed by how many times
the table enters it:

=========  ======================  ==========================================
table      NNMap 1h / 2h / 3h      NMapMap 1h / 2h / 3h
=========  ======================  ==========================================
hmf        -1 / 0 / +1             -1 / 0 / +1  (NMapMap divides by n_bar a
                                    second time, src/NMapMap_Model.cu:300,467,
                                    631; with one division it would be 0/1/2)
b_h         0 / 2 / 3               0 / 2 / 3
P_lin       0 / 1 / 2               0 / 1 / 2   (tree-level B ~ P^2)
g           1                       2           (g^2 for shear-shear)
dw/dz      -1                      -1
p_lens1     1                       1
p_lens2     1                       0           (NMapMap has one lens sample)
=========  ======================  ==========================================

Local versions pin which coordinate each look-up uses:

* b_h doubled in one mass column j: the term doubles once per halo mass whose
  nearest (round) mass index is j;
* concentration changed in one mass column j: a term changes only at points
  where a mass that enters one of its u_NFW factors has truncated index j --
  for NMapMap 3-halo that is m1 or m2 and not m3, whose occupation enters
  with u_NFW(k1, m2, f) (src/NMapMap_Model.cu:616-626);
* hmf or P_lin doubled in one redshift row: only points whose nearest
  redshift index is that row change, by the global exponent;
* HOD limit: with Ncen = 0 (central amplitude 0), satellite slope 1 and no
  NsatNsat correction, doubling M' of one population halves its Nsat and its
  n_bar exactly, so every NNMap term is unchanged and every NMapMap term
  doubles (Nsat / n_bar^2).
"""
from __future__ import annotations

import math
from dataclasses import replace

import numpy as np

from g3lhalo_amd import tables as T

ARCMIN = T.ARCMIN_IN_RAD
HOD_A = (0.3, 1.0, 1e12, 0.5, 10 ** 13.3, 1.0)
HOD_B = (0.6, 0.8, 3e11, 0.3, 2e13, 0.8)
A_CORR, EPS_CORR = 0.3, 0.2
THETA = np.array([1.7, 4.1, 2.6]) * ARCMIN
TERMS = [(0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3)]
NPTS = 1500

DEGREE = {
    "hmf": {(0, 1): -1, (0, 2): 0, (0, 3): 1, (1, 1): -1, (1, 2): 0, (1, 3): 1},
    "b_h": {(0, 1): 0, (0, 2): 2, (0, 3): 3, (1, 1): 0, (1, 2): 2, (1, 3): 3},
    "P_lin": {(0, 1): 0, (0, 2): 1, (0, 3): 2, (1, 1): 0, (1, 2): 1, (1, 3): 2},
    "g": {(0, t): 1 for t in (1, 2, 3)} | {(1, t): 2 for t in (1, 2, 3)},
    "dwdz": {(s, t): -1 for s in (0, 1) for t in (1, 2, 3)},
    "p_lens1": {(s, t): 1 for s in (0, 1) for t in (1, 2, 3)},
    "p_lens2": {(0, t): 1 for t in (1, 2, 3)} | {(1, t): 0 for t in (1, 2, 3)},
}

# mass slots (coordinate indices) of each term that enter a u_NFW factor
U_MASSES = {(0, 1): (3,), (0, 2): (3, 4), (0, 3): (3, 4, 5),
            (1, 1): (3,), (1, 2): (3, 4), (1, 3): (3, 4)}


def mass_dims(term):
    return tuple(range(3, 3 + term))


def scaled(t: T.Tables, name: str, factor: float = 2.0, row=None, col=None) -> T.Tables:
    """A copy of t with table `name` multiplied by factor (a 2-D table only in
    redshift row `row` or mass column `col` when given)."""
    a = np.array(getattr(t, name), dtype=np.float64, copy=True)
    if row is None and col is None:
        a *= factor
    else:
        a2 = a.reshape(t.nbins, t.nbins)
        if row is not None:
            a2[row, :] *= factor
        if col is not None:
            a2[:, col] *= factor
        a = a2.reshape(-1)
    return replace(t, **{name: np.ascontiguousarray(a)})


def borders(t: T.Tables):
    """Integration borders (src/NNMap_Model.cu:46-49)."""
    nb = t.nbins
    zmax_i = t.zmax - (t.zmax - t.zmin) / nb
    mmax_i = math.exp(math.log(t.mmin) + math.log(t.mmax / t.mmin) * (nb - 1) / nb)
    return t.zmin, zmax_i, t.mmin, mmax_i


def points(t: T.Tables, term: int, n: int, seed: int, mass_bin=None, mass_round=True,
           z_row=None) -> np.ndarray:
    """Random points of a term's integration box (log10 l1, log10 l2, phi,
    log10 m..., z), with l where the aperture filters are not negligible.
    mass_bin = j puts each mass coordinate, with probability 1/2, inside mass
    bin j (of the round or the truncated index); z_row = i puts half of the
    points' redshifts inside row i of the round redshift index.  Points stay
    away from index ties."""
    rng = np.random.default_rng(seed)
    zmin, zmax_i, mmin, mmax_i = borders(t)
    D = 4 + term
    x = np.empty((n, D))
    x[:, 0] = rng.uniform(0.0, 4.3, n)
    x[:, 1] = rng.uniform(0.0, 4.3, n)
    x[:, 2] = rng.uniform(0.0, 2 * 3.14159, n)
    lmr = math.log(t.mmax / t.mmin)
    for d in mass_dims(term):
        x[:, d] = rng.uniform(math.log10(mmin), math.log10(mmax_i), n)
        if mass_bin is not None:
            pick = rng.random(n) < 0.5
            u = rng.uniform(0.1, 0.9, n)
            pos = (mass_bin - 0.5 + u) if mass_round else (mass_bin + u)
            x[pick, d] = (math.log(t.mmin) + pos[pick] * lmr / t.nbins) / math.log(10)
    dz = (t.zmax - t.zmin) / t.nbins
    x[:, D - 1] = rng.uniform(zmin, zmax_i, n)
    if z_row is not None:
        pick = rng.random(n) < 0.5
        x[pick, D - 1] = t.zmin + (z_row + rng.uniform(-0.4, 0.4, pick.sum())) * dz
    return x


def mass_index(t: T.Tables, log10m: np.ndarray, rounded: bool) -> np.ndarray:
    L = np.log(10.0 ** log10m / t.mmin) * t.nbins / math.log(t.mmax / t.mmin)
    return (np.round(L) if rounded else np.floor(L)).astype(int)


def z_index(t: T.Tables, z: np.ndarray) -> np.ndarray:
    return np.round((z - t.zmin) * t.nbins / (t.zmax - t.zmin)).astype(int)


def usable(f: np.ndarray) -> np.ndarray:
    """Points whose value is a normal, finite, nonzero double (an exact power
    of two change then shows bit for bit)."""
    a = np.abs(f)
    return np.isfinite(f) & (a > 1e-280) & (a < 1e280)


# ---------------------------------------------------------------- checks
# `make(tables)` returns an evaluator e with e(stat, term, x) -> integrand
# values of the two-population HOD set (HOD_A, HOD_B, A_CORR, EPS_CORR) at
# types (1, 2); `make_hod(tables, hod_a, hod_b)` the same for other HODs
# without the NsatNsat correction.

def check_global_scaling(make, t, name):
    base, sc = make(t), make(scaled(t, name))
    for k, (stat, term) in enumerate(TERMS):
        x = points(t, term, NPTS, 100 + k)
        f0, f1 = base(stat, term, x), sc(stat, term, x)
        ok = usable(f0)
        assert ok.mean() > 0.5, (name, stat, term, ok.mean())
        want = np.ldexp(f0[ok], DEGREE[name][(stat, term)])
        bad = np.flatnonzero(f1[ok] != want)
        assert bad.size == 0, (name, stat, term, f0[ok][bad[:3]], f1[ok][bad[:3]])


def check_bh_column(make, t, col):
    base, sc = make(t), make(scaled(t, "b_h", col=col))
    for k, (stat, term) in enumerate(TERMS):
        x = points(t, term, NPTS, 200 + k, mass_bin=col, mass_round=True)
        f0, f1 = base(stat, term, x), sc(stat, term, x)
        ok = usable(f0)
        if term == 1:
            deg = np.zeros(len(x), dtype=int)
        else:
            deg = sum((mass_index(t, x[:, d], True) == col).astype(int) for d in mass_dims(term))
        assert (deg[ok] > 0).sum() > 50 or term == 1
        assert np.array_equal(f1[ok], np.ldexp(f0[ok], deg[ok])), (stat, term)


def check_conc_column(make, t, col):
    base, sc = make(t), make(scaled(t, "concentration", factor=1.37, col=col))
    for k, (stat, term) in enumerate(TERMS):
        x = points(t, term, NPTS, 300 + k, mass_bin=col, mass_round=False)
        f0, f1 = base(stat, term, x), sc(stat, term, x)
        ok = usable(f0)
        hit = np.zeros(len(x), dtype=bool)
        for d in U_MASSES[(stat, term)]:
            hit |= mass_index(t, x[:, d], False) == col
        # points whose u_NFW masses all lie outside column col: bit-identical
        assert np.array_equal(f1[ok & ~hit], f0[ok & ~hit]), (stat, term)
        # the rest move
        moved = f1[ok & hit] != f0[ok & hit]
        assert moved.size > 100 and moved.mean() > 0.9, (stat, term, moved.size, moved.mean())
        if (stat, term) == (1, 3):
            # m3 alone in column col changes nothing (Q4: u_NFW(k1, m2, f))
            only3 = ok & ~hit & (mass_index(t, x[:, 5], False) == col)
            assert only3.sum() > 50


def check_row(make, t, name, row):
    base, sc = make(t), make(scaled(t, name, row=row))
    for k, (stat, term) in enumerate(TERMS):
        x = points(t, term, NPTS, 400 + k, z_row=row)
        f0, f1 = base(stat, term, x), sc(stat, term, x)
        ok = usable(f0)
        inrow = z_index(t, x[:, -1]) == row
        assert (ok & inrow).sum() > 200
        assert np.array_equal(f1[ok & ~inrow], f0[ok & ~inrow]), (name, stat, term)
        want = np.ldexp(f0[ok & inrow], DEGREE[name][(stat, term)])
        assert np.array_equal(f1[ok & inrow], want), (name, stat, term)


def check_hod_limit(make_hod, t, rtol):
    """Ncen = 0, beta = 1, A = 0: doubling M' of population 1 leaves NNMap
    unchanged and doubles NMapMap (whose one HOD is population 1)."""
    ha = (0.3, 0.0, 1e12, 0.5, 10 ** 13.3, 1.0)
    hb = (0.6, 0.0, 3e11, 0.3, 2e13, 1.0)
    ha2 = ha[:4] + (2 * ha[4],) + ha[5:]
    base, sc = make_hod(t, ha, hb), make_hod(t, ha2, hb)
    for k, (stat, term) in enumerate(TERMS):
        x = points(t, term, NPTS, 500 + k)
        f0, f1 = base(stat, term, x), sc(stat, term, x)
        ok = usable(f0)
        assert ok.mean() > 0.5
        want = f0[ok] * (1.0 if stat == 0 else 2.0)
        np.testing.assert_allclose(f1[ok], want, rtol=rtol, atol=0)
