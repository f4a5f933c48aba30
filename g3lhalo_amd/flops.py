"""Algorithmic FP64 work per integrand evaluation, frozen per kind.

Convention: +, -, *, /, sqrt, compare, round/trunc = 1 flop; each libm call
(exp, log, pow, pow10, erf, sin, cos) = 20; sincos = 40; one Si/Ci pair =
84 (its x < 4 branch: two rational functions and a log); constant
sub-expressions (log(mmax/mmin), zmax - zmin, ...) are not counted.
Sub-expressions the reference evaluates repeatedly with equal arguments are
counted once (the device code evaluates them once).  The tallies follow
g3lhalo_amd/csrc/physics.h line by line; the reference functions are
src/NNMap_Model.cu:544-1037, src/NMapMap_Model.cu:263-632 and
src/generalFunctions.cu:6-244.

These numbers are the work credited per evaluation in integrals/s and
roofline reports: achieved = sum(evaluations x flops) / kernel time.
"""
from __future__ import annotations

POW10 = 20
APERTURE = 1 + 2 + 2 + 20 + 1          # eta, eta^2/2, -eta^2/2, exp, product
ZIDX = 4                               # (z - zmin) * nb / dz, round
MIDX = 1 + 20 + 1 + 1 + 1              # m/mmin, log, *nb, /log(..), round
KIDX = MIDX
PK = KIDX + 1                          # index + branch
OCC = 21 + 2 + 20 + 1 + 2 + (1 + 1 + 20 + 1)   # log, scale, erf, 1+, Ncen, Nsat
NSS_CROSS = 1 + 1 + 20 + 1 + 1 + 1 + 1 + 1    # Ns1Ns2 + A m^eps sqrt(..), clamp
R200 = ZIDX + 3 + 20                   # pow(0.239 m / rho / 200, 1/3)
NFW_MASS = ZIDX + MIDX                 # concentration look-up indices per (m, z)
NFW_F = 1 + 1 + 20 + 1 + 1 + 1         # c = f conc; log(1+c) - c/(1+c)
SICI = 84
UNFW = 2 + 2 + 2 * SICI + 40 + 2 + 2 + (1 + 20 + 2) + 3
L3 = 8                                 # sqrt(l1^2 + l2^2 + 2 l1 l2 cos)
KS = 3
FK = 9                                 # F(k1, k2, c)
BILIN = 8 + 8 + 8 + ZIDX + 3 * PK + 3 * FK + 9

FLOPS_PER_EVAL = {
    # NNMap 1-halo: 3 pow10, cos, indices, l3, k, r200, 2 HODs, NsatNsat,
    # 3 concentrations, 3 u_NFW, bispectrum and projection products, 3 filters
    1: 3 * POW10 + 20 + ZIDX + MIDX + L3 + KS + R200 + 2 * OCC + NSS_CROSS + NFW_MASS
       + 3 * NFW_F + 3 * UNFW + 7 + 8 + 1 + 9 + 8 + 3 * APERTURE,
    # NNMap 2-halo: 4 pow10, 3 P(k), 2 r200, 4 HOD evaluations, 6 u_NFW
    2: 4 * POW10 + 20 + ZIDX + 2 * MIDX + L3 + KS + 3 * PK + 2 * R200 + 4 * OCC + NSS_CROSS
       + 2 * NFW_MASS + 6 * NFW_F + 6 * UNFW + 8 + 8 + 22 + 10 + 9 + 3 * APERTURE,
    # NNMap 3-halo: 5 pow10, 3 r200, 3 u_NFW, 2 HODs, tree-level bispectrum
    3: 5 * POW10 + 20 + ZIDX + 3 * MIDX + L3 + KS + 3 * R200 + 3 * NFW_MASS + 3 * NFW_F
       + 3 * UNFW + 2 * OCC + 4 + BILIN + 15 + 9 + 10 + 3 * APERTURE,
    # NMapMap 1-halo
    4: 3 * POW10 + 20 + ZIDX + MIDX + L3 + KS + OCC + R200 + NFW_MASS + 2 * NFW_F
       + 3 * UNFW + 11 + 10 + 9 + 3 * APERTURE,
    # NMapMap 2-halo
    5: 4 * POW10 + 20 + ZIDX + 2 * MIDX + L3 + KS + 3 * PK + 2 * OCC + 2 * R200
       + 2 * NFW_MASS + 4 * NFW_F + 6 * UNFW + 30 + 10 + 10 + 3 * APERTURE,
    # NMapMap 3-halo
    6: 5 * POW10 + 20 + ZIDX + 3 * MIDX + L3 + KS + OCC + 2 * R200 + 2 * NFW_MASS
       + 3 * NFW_F + 3 * UNFW + BILIN + 17 + 10 + 11 + 3 * APERTURE,
    # n_bar: pow10, one HOD, two indices, m hmf (Nc + Ns)
    0: POW10 + OCC + MIDX + ZIDX + 3,
}


def kind_of(stat: int, term: int) -> int:
    return term if stat == 0 else 3 + term


def flops_per_eval(stat: int, term: int) -> int:
    return FLOPS_PER_EVAL[kind_of(stat, term)]


# FP32 share of the algorithmic work of the mixed-precision integrand at
# most: the u_NFW core (all but the three FP64 argument products; FP32 where
# k r_s is not small, FP64 below) and the exponential of each aperture
# filter.  Everything else (indices, masses, r200, concentrations, HOD
# occupations, table values, products) stays FP64.
FP32_PER_EVAL = {
    1: 3 * (UNFW - 3) + 3 * 20,
    2: 6 * (UNFW - 3) + 3 * 20,
    3: 3 * (UNFW - 3) + 3 * 20,
    4: 3 * (UNFW - 3) + 3 * 20,
    5: 6 * (UNFW - 3) + 3 * 20,
    6: 3 * (UNFW - 3) + 3 * 20,
    0: 0,
}

FP64_PEAK = 78.6    # TFLOP/s, MI355X vector FP64
FP32_PEAK = 157.3   # TFLOP/s, MI355X vector FP32 (packed / dual-issue instructions)
# The FP32 rate of the instructions this build emits: packed FP32 is switched
# off (-packed-fp32-ops), so an FP32 FMA issues one per lane per cycle, as an
# FP64 FMA does: 256 CUs x 4 SIMDs x 16 lanes x 2 flops x 2.4 GHz.
FP32_UNPACKED_PEAK = 78.6


def fp64_fraction(kind: int) -> float:
    return 1.0 - FP32_PER_EVAL[kind] / FLOPS_PER_EVAL[kind]


def mixed_peak_tflops(kind: int, packed: bool = False) -> float:
    """1 / (phi64 / FP64 peak + (1 - phi64) / FP32 peak), with the FP32 peak of
    the instructions emitted (unpacked unless packed=True)."""
    p = fp64_fraction(kind)
    f32 = FP32_PEAK if packed else FP32_UNPACKED_PEAK
    return 1.0 / (p / FP64_PEAK + (1.0 - p) / f32)
