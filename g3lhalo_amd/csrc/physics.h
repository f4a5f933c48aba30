// physics.h -- device integrands of the NNMap / NMapMap path (gfx950).
//
// The six term integrands and the n_bar integrand are written once, as
// templates over a precision policy:
//   P64    FP64 throughout, following the reference's expression order
//          (src/generalFunctions.cu:6-147, src/HOD.cu:35-94,
//           src/NNMap_Model.cu:544-1108, src/NMapMap_Model.cu:263-660).
//   PMix   FP32 for the transcendental bulk (the u_NFW core with its Si/Ci
//          where k r_s is not small, the exponentials of the aperture
//          filters), FP64 for the small-k r_s branch of u_NFW, the HOD
//          occupations (erf and the log-space mass exponentials) and
//          everything that decides a table index or carries a large dynamic
//          range (coordinates, masses, the nearest-index chains, table
//          values, products, the NFW normalisation).
//
// Deviations of P64 from the reference that change only the last bits:
// u_NFW by the cancellation-free branch identities noted at u_nfw, with
// Si/Ci's polynomials summed over precomputed powers, sin/cos by our own
// Cody-Waite/Taylor routine and the ln x of Ci(x) by fdlibm's method
// (log_pos); the HOD logarithms and powers from log(m/mmin); cbrt for
// pow(., 1/3); exp10 for the 10^x of l1, l2; the redshift-only projection
// factor multiplied into B in one step; shared sin/cos of one argument;
// sub-expressions the reference evaluates repeatedly with equal arguments
// evaluated once.  The NNMap 3-halo term,
// whose integrals are sensitive to last-bit differences, keeps the
// reference's HOD forms (P64R) and its order for the projection factor.
// Masses keep pow(10, x) and the index chains keep the reference's exact
// order: region centres sit exactly on rounding ties of the nearest-index
// look-ups (the root centre maps to mass index 63.5 at 128 bins).
#pragma once
#include <hip/hip_runtime.h>

#include "factor.h"
#include "layout.h"

#pragma clang fp contract(off)

// Device functions, also compiled for the host by accuracy tools.
#ifndef G3L_HD
#define G3L_HD __host__ __device__ __forceinline__
#endif

namespace g3l {

// a / b for the normal, finite operands of the u_NFW arithmetic: the hardware
// reciprocal estimate refined by two Newton-Raphson steps, then one residual
// correction of the quotient (within one unit in the last place; about a third
// of the instructions of the IEEE division sequence, which also handles
// denormal and infinite operands that cannot occur here).
G3L_HD double fdiv(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  r = fma(fma(-b, r, 1.0), r, r);
  r = fma(fma(-b, r, 1.0), r, r);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
#else
  return a / b;
#endif
}

// Natural logarithm of a positive, finite, normal x, by the method of fdlibm's
// e_log.c (Sun Microsystems, 1993): x = 2^k (1 + f) with 1 + f in
// [sqrt(2)/2, sqrt(2)), s = f / (2 + f), log(1 + f) = 2 atanh(s)
// = f - (f^2/2 - s (f^2/2 + R)) with R(s^2) the degree-14 minimax polynomial
// of that method, and k ln 2 added in two parts; error below one unit in the
// last place.  About 40 % of the library's instructions, which also handle
// zero, negative, infinite and denormal arguments.  Used for the Ci(x) of
// u_NFW, whose arguments are in (0, 4); the look-up indices keep the
// library's log.
G3L_HD double log_pos(double x) {
  int k;
  double m = frexp(x, &k);   // x = m 2^k, m in [1/2, 1)
  if (m < 0.70710678118654752440) {
    m += m;
    k -= 1;
  }
  const double f = m - 1.0;
  const double s = fdiv(f, 2.0 + f);
  const double z = s * s, w = z * z;
  const double t1 = w * fma(w, fma(w, 1.531383769920937332e-01, 2.222219843214978396e-01),
                            3.999999999940941908e-01);
  const double t2 =
      z * fma(w, fma(w, fma(w, 1.479819860511658591e-01, 1.818357216161805012e-01),
                     2.857142874366239149e-01),
              6.666666666666735130e-01);
  const double R = t2 + t1;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  return dk * 6.93147180369123816490e-01 -
         ((hfsq - (s * (hfsq + R) + dk * 1.90821492927058770002e-10)) - f);
}

// sin and cos together for |x| < 2^20; larger arguments take the library's
// sincos.  x = n pi/2 + r by Cody and Waite's additive reduction (Cody &
// Waite, "Software Manual for the Elementary Functions", 1980) in three FMA
// steps (pi/2 split into
// three doubles, so n pi/2 is carried to ~1e-49 relative and r has an
// absolute error near 1e-16), |r| <= pi/4 up to rounding; then the Taylor
// series of sin r through r^17 and of cos r through r^16, whose truncation
// errors at |r| = pi/4 are below 1e-19 and 3e-18, summed over precomputed
// powers of r^2 (coefficients as FMA multiplicands, in scalar registers);
// n mod 4 selects and signs the pair.  The NFW arguments k r_s stay far below
// 2^20 in every workload here.
// x = n pi/2 + r, |r| <= pi/4 (up to rounding), q = n.
G3L_HD double sincos_reduce(double x, int &q) {
  const double n = rint(x * 0.6366197723675814);
  double r = fma(-n, 1.5707963267948966, x);
  r = fma(-n, 6.123233995736766e-17, r);
  r = fma(-n, -1.4973849048591698e-33, r);
  q = (int)n;
  return r;
}
// sin x and cos x from sin r, cos r and q = n mod 4.
G3L_HD void sincos_select(int q, double sr, double cr, double &s, double &c) {
  const double a = (q & 1) ? cr : sr;
  const double b = (q & 1) ? sr : cr;
  s = (q & 2) ? -a : a;
  c = ((q + 1) & 2) ? -b : b;
}
G3L_HD void sincos_cw(double x, double &s, double &c) {
  if (!(fabs(x) < 1048576.0)) {
    sincos(x, &s, &c);
    return;
  }
  int q;
  const double r = sincos_reduce(x, q);
  const double z = r * r;
  const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
  const double z5 = z4 * z, z6 = z3 * z3, z7 = z4 * z3, z8 = z4 * z4;
  double ps = -0.16666666666666666 * z;
  ps = fma(0.008333333333333333, z2, ps);
  ps = fma(-0.0001984126984126984, z3, ps);
  ps = fma(2.7557319223985893e-06, z4, ps);
  ps = fma(-2.505210838544172e-08, z5, ps);
  ps = fma(1.6059043836821613e-10, z6, ps);
  ps = fma(-7.647163731819816e-13, z7, ps);
  ps = fma(2.8114572543455206e-15, z8, ps);
  double pc = -0.5 * z;
  pc = fma(0.041666666666666664, z2, pc);
  pc = fma(-0.001388888888888889, z3, pc);
  pc = fma(2.48015873015873e-05, z4, pc);
  pc = fma(-2.755731922398589e-07, z5, pc);
  pc = fma(2.08767569878681e-09, z6, pc);
  pc = fma(-1.1470745597729725e-11, z7, pc);
  pc = fma(4.779477332387385e-14, z8, pc);
  sincos_select(q, fma(r, ps, r), 1.0 + pc, s, c);
}
// sin and cos together as sincos_cw, with the minimax polynomials of
// fdlibm's __kernel_sin and __kernel_cos (Sun Microsystems, 1993: degree 13
// and 14 in r on |r| <= pi/4, approximation errors below 4e-18 relative)
// instead of the Taylor series: two terms and one power of r^2 fewer.
G3L_HD void sincos_mm(double x, double &s, double &c) {
  if (!(fabs(x) < 1048576.0)) {
    sincos(x, &s, &c);
    return;
  }
  int q;
  const double r = sincos_reduce(x, q);
  const double z = r * r;
  const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
  const double z5 = z4 * z, z6 = z3 * z3, z7 = z4 * z3;
  double ps = -1.66666666666666324348e-01 * z;
  ps = fma(8.33333333332248946124e-03, z2, ps);
  ps = fma(-1.98412698298579493134e-04, z3, ps);
  ps = fma(2.75573137070700676789e-06, z4, ps);
  ps = fma(-2.50507602534068634195e-08, z5, ps);
  ps = fma(1.58969099521155010221e-10, z6, ps);
  double pc = -0.5 * z;
  pc = fma(4.16666666666666019037e-02, z2, pc);
  pc = fma(-1.38888888888741095749e-03, z3, pc);
  pc = fma(2.48015872894767294178e-05, z4, pc);
  pc = fma(-2.75573143513906633035e-07, z5, pc);
  pc = fma(2.08757232129817482790e-09, z6, pc);
  pc = fma(-1.13596475577881948265e-11, z7, pc);
  sincos_select(q, fma(r, ps, r), 1.0 + pc, s, c);
}

// Tables resident in HBM (z-major 2-D tables, nb*nb + nb for P_lin).
struct DevTables {
  int nb;
  double zmin, zmax, kmin, kmax, mmin, mmax;
  double lmr;    // log(mmax / mmin), computed on the host
  double lkr;    // log(kmax / kmin), computed on the host
  const double *g, *p1, *p2, *w, *dwdz, *hmf, *plin, *bh, *conc, *rho;
};

struct DevHod {
  double f, alpha, mth, sigma, mprime, beta;
};

// Per-integral parameters seen by the integrand.
struct TermArgs {
  double t1, t2, t3;        // apertures [rad]
  DevHod h1, h2;            // HODs selected by type1, type2 (NMapMap: h1)
  const double *nb1, *nb2;  // n_bar selected by type (NMapMap: nb1)
  double A, eps;
  int same;                 // NsatNsat same-type branch
  // log-space HOD constants of PMix: log(mmin/M_th), log(mmin/M'), log(mmin)
  double lth1, lmp1, lth2, lmp2, lmmin;
  // 1-D integrands: kOneDNbar, or a table-precomputation integral whose
  // cosmology and transfer-function parameters are at aux (PrecompParams)
  int mode;
  const double *aux;
};

// ---------------------------------------------------- table precomputation
// The 1-D integrands of the reference's table precomputation
// (src/Cosmology.cpp:84-118, src/Powerspectrum.cpp:71-284), evaluated by the
// batched GK15 engine like n_bar, with the per-integral parameters in t1..t3.
enum OneDMode : int {
  kOneDNbar = 0,       // n_bar(z) (t1 = z)
  kOneDComoving = 1,   // c / H0 / sqrt(E(z))               -> w(z)
  kOneDAmplitude = 2,  // W(kR)^2 T(k)^2 k^(ns+2), t1 = R    -> sigma8 amplitude
  kOneDSigma2 = 3,     // k^2 W(kR)^2 P(k, z), t1 = R, t2 = D(z)/D(0)
  kOneDDSigma2 = 4,    // k^2 W W'(kR) dR P(k, z), t1 = R, t2 = D(z)/D(0), t3 = dR
};

// Eisenstein & Hu transfer-function parameters and the cosmology, set on the
// host (src/Powerspectrum.cpp:15-69, src/Cosmology.cpp:11-72).
struct PrecompParams {
  double Om, OL, H0, c, ns, amplitude;
  double omhh, obhh, k_equality, sound_horizon, k_silk, alpha_c, beta_c, alpha_b, beta_node,
      beta_b;
};

// T(k) (src/Powerspectrum.cpp:160-197), in the reference's evaluation order.
G3L_HD double eh_transfer(const PrecompParams &p, double k) {
  if (k == 0.0) return 1.0;
  const double q = k / 13.41 / p.k_equality;
  const double xx = k * p.sound_horizon;
  const double T_c_ln_beta = log(2.718282 + 1.8 * p.beta_c * q);
  const double T_c_ln_nobeta = log(2.718282 + 1.8 * q);
  const double pq = pow(q, 1.08);
  const double T_c_C_alpha = 14.2 / p.alpha_c + 386.0 / (1.0 + 69.9 * pq);
  const double T_c_C_noalpha = 14.2 + 386.0 / (1 + 69.9 * pq);
  const double T_c_f = 1.0 / (1.0 + pow(xx / 5.4, 4));
  const double T_c = T_c_f * T_c_ln_beta / (T_c_ln_beta + T_c_C_noalpha * q * q) +
                     (1. - T_c_f) * T_c_ln_beta / (T_c_ln_beta + T_c_C_alpha * q * q);
  const double s_tilde = p.sound_horizon * pow(1. + pow(p.beta_node / xx, 3), -1. / 3.);
  const double xx_tilde = k * s_tilde;
  const double T_b_T0 = T_c_ln_nobeta / (T_c_ln_nobeta + T_c_C_noalpha * q * q);
  const double T_b = sin(xx_tilde) / (xx_tilde) *
                     (T_b_T0 / (1. + pow(xx / 5.2, 2)) +
                      p.alpha_b / (1. + pow(p.beta_b / xx, 3)) * exp(-pow(k / p.k_silk, 1.4)));
  const double f_baryon = p.obhh / p.omhh;
  return f_baryon * T_b + (1. - f_baryon) * T_c;
}

// Top-hat window and its derivative (src/Powerspectrum.cpp:97-107).
G3L_HD double tophat(double k, double R) {
  const double x = k * R;
  return 3 / x / x / x * (sin(x) - x * cos(x));
}
G3L_HD double d_tophat(double k, double R) {
  const double x = k * R;
  return ((x * x - 3.) * sin(x) + 3. * x * cos(x)) * 3 * k / x / x / x / x;
}

// Linear P(k, z) with growth = D(z)/D(0) (src/Powerspectrum.cpp:200-207).
G3L_HD double eh_spectrum(const PrecompParams &p, double k, double growth) {
  const double trans = eh_transfer(p, k);
  return p.amplitude * trans * trans * growth * growth * pow(k, p.ns);
}

// One precomputation integrand value at x (src/Cosmology.cpp:104-118,
// src/Powerspectrum.cpp:71-95, 229-245, 265-284).
G3L_HD double precomp_kernel(const TermArgs &a, double x) {
  const PrecompParams &p = *reinterpret_cast<const PrecompParams *>(a.aux);
  switch (a.mode) {
    case kOneDComoving: {
      const double zp = 1 + x;
      return p.c / p.H0 / sqrt(p.Om * zp * zp * zp + p.OL);
    }
    case kOneDAmplitude: {
      const double W = tophat(x, a.t1);
      const double T = eh_transfer(p, x);
      return W * W * T * T * pow(x, p.ns + 2);
    }
    case kOneDSigma2: {
      const double W = tophat(x, a.t1);
      return x * x * W * W * eh_spectrum(p, x, a.t2);
    }
    default: {
      const double W = tophat(x, a.t1);
      const double Wprime = d_tophat(x, a.t1);
      return x * x * W * Wprime * a.t3 * eh_spectrum(p, x, a.t2);
    }
  }
}

// ---------------------------------------------------------------- indices
// Nearest-index look-ups: round() in the integrands, int() truncation in
// u_NFW and r_200 (src/generalFunctions.cu:11-12,35).  Every index is clamped
// to the allocated table; inside the integration box the clamp never acts,
// it only keeps invalid input (e.g. a zero comoving distance) from reading
// outside the tables.
G3L_HD int clampi(int i, int hi) { return i < 0 ? 0 : (i > hi ? hi : i); }
G3L_HD int zix_round(const DevTables &T, double z) {
  return clampi((int)round((z - T.zmin) * T.nb / (T.zmax - T.zmin)), T.nb - 1);
}
G3L_HD int zix_trunc(const DevTables &T, double z) {
  return clampi((int)((z - T.zmin) * T.nb / (T.zmax - T.zmin)), T.nb - 1);
}
// Held in [-1, nb^2 + nb] before the conversion: an infinite k (a zero
// comoving distance) would otherwise convert out of int's range, and
// zi * nb + ki could overflow.  Every finite k in that range keeps its
// index; beyond it the row-spill read lands past the padded table either
// way and pk's clamp gives the same entry.  NaN maps to 0.
G3L_HD int kix_round(const DevTables &T, double k) {
  const double x = round(log(k / T.kmin) * T.nb / T.lkr);
  const double top = (double)T.nb * T.nb + T.nb;
  return x >= 0 ? (int)fmin(x, top) : (x < 0 ? (int)fmax(x, -1.0) : 0);
}
// P(k) look-up, linear below kmin, no upper guard (src/NNMap_Model.cu:819-839):
// an index past the row reads on into the next row; P_lin carries nb
// trailing zeros for the last row.
G3L_HD double pk(const DevTables &T, int zi, double k) {
  int ki = kix_round(T, k);
  if (ki >= 0) return T.plin[clampi(zi * T.nb + ki, T.nb * T.nb + T.nb - 1)];
  return T.plin[zi * T.nb] / T.kmin * k;
}

// One halo mass at one redshift: m = 10^x with the reference's rounding, its
// look-up indices and r_200 (src/generalFunctions.cu:32-37).
struct Mass {
  double m, L;   // L = log(m / mmin)
  int ir, it;    // round and int() indices
  double r200;
};
template <class P>
G3L_HD Mass make_mass(const DevTables &T, double x, int zt) {
  Mass M;
  M.m = pow(10.0, x);
  M.L = log(M.m / T.mmin);
  M.ir = clampi((int)round(M.L * T.nb / T.lmr), T.nb - 1);
  M.it = clampi((int)(M.L * T.nb / T.lmr), T.nb - 1);
  M.r200 = P::cbrt3(0.239 * M.m / T.rho[zt] / 200);
  return M;
}

// NFW normalisation for one (mass, f): c = f conc[z, m] with int() indices,
// F = ln(1+c) - c/(1+c) (src/generalFunctions.cu:14-27).
struct Nfw {
  double r200, c, lc, F;   // lc = ln(1+c)
};
G3L_HD Nfw make_nfw(const DevTables &T, const Mass &M, int zt, double f) {
  Nfw n;
  n.r200 = M.r200;
  n.c = f * T.conc[zt * T.nb + M.it];
  n.lc = log(1. + n.c);
  n.F = n.lc - n.c / (1. + n.c);
  return n;
}

struct Occ {
  double nc, ns;
};

// ----------------------------------------------------------------- Si, Ci
// Rational approximations of src/generalFunctions.cu:44-87.
//
// The polynomials are summed term by term over precomputed powers,
// acc = fma(c_k, t^k, acc), rather than by Horner's rule.  In Horner's form
// every coefficient is an addend, and this target's two-address FP64 FMA
// needs an addend in vector registers (two v_mov per coefficient); as a
// multiplicand a coefficient stays in scalar registers, so each term costs
// one vector instruction.  The powers are built by a balanced product tree
// (at most four roundings for t^10); every sum below has terms of one sign
// or a negligible opposite-sign tail, so the result stays within a few
// units in the last place of the reference's left-to-right form.

// Si(x) = sn / sd and the rational part of Ci(x) = gamma + ln x + cn / cd
// for x < 4, as numerators and denominators (sd, cd >= 1).
G3L_HD void sici_small_nd(double x, double &sn, double &sd, double &cn, double &cd) {
  const double z = x * x;
  const double z2 = z * z, z3 = z2 * z, z4 = z2 * z2;
  const double z5 = z4 * z, z6 = z3 * z3, z7 = z4 * z3;
  double a = fma(-4.54393409816329991e-2, z, 1.0);
  a = fma(1.15457225751016682e-3, z2, a);
  a = fma(-1.41018536821330254e-5, z3, a);
  a = fma(9.43280809438713025e-8, z4, a);
  a = fma(-3.53201978997168357e-10, z5, a);
  a = fma(7.08240282274875911e-13, z6, a);
  a = fma(-6.05338212010422477e-16, z7, a);
  double b = fma(1.01162145739225565e-2, z, 1.0);
  b = fma(4.99175116169755106e-5, z2, b);
  b = fma(1.55654986308745614e-7, z3, b);
  b = fma(3.28067571055789734e-10, z4, b);
  b = fma(4.5049097575386581e-13, z5, b);
  b = fma(3.21107051193712168e-16, z6, b);
  sn = x * a;
  sd = b;
  // x^2 times Ci's numerator, -x^2/4 + ..., summed from its first term
  double n = -0.25 * z;
  n = fma(7.51851524438898291e-3, z2, n);
  n = fma(-1.27528342240267686e-4, z3, n);
  n = fma(1.05297363846239184e-6, z4, n);
  n = fma(-4.68889508144848019e-9, z5, n);
  n = fma(1.06480802891189243e-11, z6, n);
  n = fma(-9.93728488857585407e-15, z7, n);
  double d = fma(1.1592605689110735e-2, z, 1.0);
  d = fma(6.72126800814254432e-5, z2, d);
  d = fma(2.55533277086129636e-7, z3, d);
  d = fma(6.97071295760958946e-10, z4, d);
  d = fma(1.38536352772778619e-12, z5, d);
  d = fma(1.89106054713059759e-15, z6, d);
  d = fma(1.39759616731376855e-18, z7, d);
  cn = n;
  cd = d;
}
// Si(x) and the rational part of Ci(x) = gamma + ln x + cr for x < 4.
G3L_HD void sici_small_parts(double x, double &si, double &cr) {
  double sn, sd, cn, cd;
  sici_small_nd(x, sn, sd, cn, cd);
  si = fdiv(sn, sd);
  cr = fdiv(cn, cd);
}

// g(x) = gn / gd and f(x) - 1/x = fn / fd for x >= 4, as numerators and
// denominators (gd >= 1, fd >= x), the latter from the exact coefficient
// differences of f's numerator and denominator (no cancellation); t = 1/x^2.
G3L_HD void sici_aux_nd(double x, double &gn, double &gd, double &fn, double &fd) {
  const double t = fdiv(1.0, x * x);
  const double t2 = t * t, t3 = t2 * t, t4 = t2 * t2, t5 = t4 * t;
  const double t6 = t3 * t3, t7 = t4 * t3, t8 = t4 * t4, t9 = t5 * t4, t10 = t5 * t5;
  // fq = 1 + O(t): f's denominator
  double fq = fma(7.46437068161927678031e2, t, 1.0);
  fq = fma(1.97865247031583951450e5, t2, fq);
  fq = fma(2.41535670165126845144e7, t3, fq);
  fq = fma(1.47478952192985464958e9, t4, fq);
  fq = fma(4.58595115847765779830e10, t5, fq);
  fq = fma(7.08501308149515401563e11, t6, fq);
  fq = fma(5.06084464593475076774e12, t7, fq);
  fq = fma(1.43468549171581016479e13, t8, fq);
  fq = fma(1.11535493509914254097e13, t9, fq);
  // t dn(t): numerator minus denominator of f, times t
  double dn = -1.99999999999097743952e0 * t;
  dn = fma(-1.46887413643708168820e3, t2, dn);
  dn = fma(-3.78536003969501121901e5, t3, dn);
  dn = fma(-4.40554837171082794666e7, t4, dn);
  dn = fma(-2.48588769773332548141e9, t5, dn);
  dn = fma(-6.79674775754933776855e10, t6, dn);
  dn = fma(-8.51162840223981323242e11, t7, dn);
  dn = fma(-4.26733661912124414062e12, t8, dn);
  dn = fma(-6.20538246899190625000e12, t9, dn);
  dn = fma(-4.94701168645415954590e11, t10, dn);
  // gp, gq = 1 + O(t): g's numerator and denominator
  double gp = fma(8.1359520115168615e2, t, 1.0);
  gp = fma(2.35239181626478200e5, t2, gp);
  gp = fma(3.12557570795778731e7, t3, gp);
  gp = fma(2.06297595146763354e9, t4, gp);
  gp = fma(6.83052205423625007e10, t5, gp);
  gp = fma(1.09049528450362786e12, t6, gp);
  gp = fma(7.57664583257834349e12, t7, gp);
  gp = fma(1.81004487464664575e13, t8, gp);
  gp = fma(6.43291613143049485e12, t9, gp);
  gp = fma(-1.36517137670871689e12, t10, gp);
  double gq = fma(8.19595201151451564e2, t, 1.0);
  gq = fma(2.40036752835578777e5, t2, gq);
  gq = fma(3.26026661647090822e7, t3, gq);
  gq = fma(2.23355543278099360e9, t4, gq);
  gq = fma(7.87465017341829930e10, t5, gq);
  gq = fma(1.39866710696414565e12, t6, gq);
  gq = fma(1.17164723371736605e13, t7, gq);
  gq = fma(4.01839087307656620e13, t8, gq);
  gq = fma(3.99653257887490811e13, t9, gq);
  gn = gp * t;
  gd = gq;
  fn = dn;
  fd = fq * x;
}
// g(x) and f(x) - 1/x for x >= 4.
G3L_HD void sici_aux_gfm(double x, double &g, double &fm) {
  double gn, gd, fn, fd;
  sici_aux_nd(x, gn, gd, fn, fd);
  g = fdiv(gn, gd);
  fm = fdiv(fn, fd);
}

// ======================================================================
// Truncated-NFW transform (src/generalFunctions.cu:6-30),
//   u F = sin a1 (Si a2 - Si a1) + cos a1 (Ci a2 - Ci a1) - sin(a1 c) / a2,
// a1 = k r200 / c, a2 = a1 (1 + c), F = ln(1+c) - c/(1+c), rearranged
// (equal in exact arithmetic) by the branch of the Si/Ci approximation
// each argument falls in:
//   a1 >= 4:       u F = g(a1) + sin(a1 c) (f(a2) - 1/a2) - g(a2) cos(a1 c);
//   a1 < 4 <= a2:  u F = pi/2 sin a1 - sin a1 Si(a1) - cos a1 Ci(a1)
//                        + sin(a1 c) (f(a2) - 1/a2) - g(a2) cos(a1 c);
//   a2 < 4:        u = 1 + [(cos a1 - 1) ln(1+c) + (a1 c - sin a1 c) / a2
//                        + cos a1 (Cr(a2) - Cr(a1)) + sin a1 (Si a2 - Si a1)] / F,
// with Ci = gamma + ln x + Cr below 4.  Every branch uses sin/cos of a1 c
// (a2 - a1) instead of a2, none has the reference's O(1) cancellations,
// and the pieces two branches share are computed once, so a wave whose
// lanes fall in different branches runs each piece at most once.
//
// u_nfw_a1 takes a1 itself (for callers that computed it to sort items by
// branch); u_nfw computes it first, by the same expression.
template <class P>
G3L_HD double u_nfw_a1(double a1, const Nfw &n) {
  const double a2 = a1 * (1 + n.c);
  double sx, cx;
  sincos_cw(a1 * n.c, sx, cx);
  double s1 = 0, c1 = 0, si1 = 0, cr1 = 0, g2 = 0, fm2 = 0;
  if (a1 < 4) {
    sincos_cw(a1, s1, c1);
    sici_small_parts(a1, si1, cr1);
  }
  if (a2 >= 4) sici_aux_gfm(a2, g2, fm2);
  if (a1 >= 4) {
    double g1, fm1;
    sici_aux_gfm(a1, g1, fm1);
    return fdiv(g1 + sx * fm2 - g2 * cx, n.F);
  }
  if (a2 >= 4) {
    const double ci1 = 0.5772156649 + P::ln_ci(a1) + cr1;
    return fdiv(0.5 * 3.14159265358979323846 * s1 - s1 * si1 - c1 * ci1 + sx * fm2 - g2 * cx,
                n.F);
  }
  double si2, cr2;
  sici_small_parts(a2, si2, cr2);
  const double rest =
      (c1 - 1) * n.lc + fdiv(a1 * n.c - sx, a2) + c1 * (cr2 - cr1) + s1 * (si2 - si1);
  return 1 + fdiv(rest, n.F);
}
// u_nfw_a1 with sin and cos by sincos_mm and each branch's quotients over
// one common denominator, so that one division replaces six (a2 < 4: Si and
// Cr at a1 and a2, (a1 c - sin a1 c) / a2 and / F), five (a1 < 4 <= a2: Si and Cr at a1, g
// and f - 1/x at a2, / F) or four (a1 >= 4: g at a1, g and f - 1/x at a2,
// / F).  Equal to u_nfw_a1 in exact arithmetic; the denominators are
// products of the approximants' denominators (each at least 1, and at most
// about 6e5 for arguments near 4 or x itself for large ones) with a2 and F,
// so nothing overflows and no new cancellation appears.  With a2 < 4, over D = a2 Cd1 Cd2 Sd1 Sd2 F:
//   u = 1 + { [a2 (cos a1 - 1) ln(1+c) + a1 c - sin a1 c] Cd1 Cd2 Sd1 Sd2
//             + a2 [cos a1 (Cn2 Cd1 - Cn1 Cd2) Sd1 Sd2
//                   + sin a1 (Sn2 Sd1 - Sn1 Sd2) Cd1 Cd2] } / D.
template <class P>
G3L_HD double u_nfw_a1_1q(double a1, const Nfw &n) {
  const double a2 = a1 * (1 + n.c);
  double sx, cx;
  sincos_mm(a1 * n.c, sx, cx);
  double s1 = 0, c1 = 0, sn1 = 0, sd1 = 1, cn1 = 0, cd1 = 1;
  double gn2 = 0, gd2 = 1, fn2 = 0, fd2 = 1;
  if (a1 < 4) {
    sincos_mm(a1, s1, c1);
    sici_small_nd(a1, sn1, sd1, cn1, cd1);
  }
  if (a2 >= 4) sici_aux_nd(a2, gn2, gd2, fn2, fd2);
  if (a1 >= 4) {
    // u F = g(a1) + sin(a1 c) (f(a2) - 1/a2) - g(a2) cos(a1 c)
    double gn1, gd1, fn1, fd1;
    sici_aux_nd(a1, gn1, gd1, fn1, fd1);
    const double p2 = fd2 * gd2;
    const double r2 = sx * fn2 * gd2 - cx * gn2 * fd2;
    return fdiv(fma(gn1, p2, gd1 * r2), gd1 * p2 * n.F);
  }
  if (a2 >= 4) {
    // u F = pi/2 sin a1 - cos a1 (gamma + ln a1) - sin a1 Si(a1) - cos a1 Cr(a1)
    //       + sin(a1 c) (f(a2) - 1/a2) - g(a2) cos(a1 c)
    const double A = 0.5 * 3.14159265358979323846 * s1 - c1 * (0.5772156649 + P::ln_ci(a1));
    const double p1 = sd1 * cd1, p2 = fd2 * gd2;
    const double r1 = s1 * sn1 * cd1 + c1 * cn1 * sd1;
    const double r2 = sx * fn2 * gd2 - cx * gn2 * fd2;
    return fdiv(fma(A, p1, -r1) * p2 + r2 * p1, p1 * p2 * n.F);
  }
  double sn2, sd2, cn2, cd2;
  sici_small_nd(a2, sn2, sd2, cn2, cd2);
  const double C = cd1 * cd2, S = sd1 * sd2, CS = C * S;
  const double dC = fma(cn2, cd1, -cn1 * cd2);
  const double dS = fma(sn2, sd1, -sn1 * sd2);
  const double t0 = fma((c1 - 1) * n.lc, a2, a1 * n.c - sx);
  const double num = fma(t0, CS, a2 * fma(c1 * dC, S, s1 * dS * C));
  return 1 + fdiv(num, a2 * CS * n.F);
}
// a1 = k r200 / c of u_NFW.
G3L_HD double nfw_a1(const Nfw &n, double k) { return fdiv(k * n.r200, n.c); }
template <class P>
G3L_HD double u_nfw(const Nfw &n, double k) {
  return u_nfw_a1<P>(nfw_a1(n, k), n);
}

// FP64 policy
struct P64 {
  static constexpr bool kMixed = false;
  G3L_HD static double l(double x) { return exp10(x); }
  // x^(1/3) of r_200
  G3L_HD static double cbrt3(double x) { return cbrt(x); }

  // Ncen and Nsat share log(m/M_th)/sigma/sqrt2 and its erf (src/HOD.cu:35-76);
  // the logarithms come from L = log(m/mmin): log(m/M_th) = L + log(mmin/M_th),
  // (m/M')^beta = exp(beta (L + log(mmin/M'))).
  G3L_HD static Occ occ(const DevHod &h, double lth, double lmp, const Mass &M) {
    double e = 1 + erf((M.L + lth) / h.sigma / 1.414213562);
    Occ o;
    o.nc = 0.5 * h.alpha * e;
    o.ns = 0.5 * e * exp(h.beta * (M.L + lmp));
    return o;
  }

  // m^eps of NsatNsat's cross term, exp(eps (L + log mmin)).
  G3L_HD static double mpow(const Mass &M, double eps, double lmmin) {
    return exp(eps * (M.L + lmmin));
  }

  // Truncated-NFW transform (u_nfw above); ln x of Ci(x) by log_pos.
  G3L_HD static double ln_ci(double x) { return log_pos(x); }
  G3L_HD static double u(const Nfw &n, double k);
  G3L_HD static double u_a1(double a1, const Nfw &n);

  // U(eta) = eta^2/2 exp(-eta^2/2) (src/generalFunctions.cu:39-42).
  G3L_HD static double ap(double eta) {
    return 0.5 * eta * eta * exp(-0.5 * eta * eta);
  }
};

G3L_HD double P64::u(const Nfw &n, double k) { return u_nfw_a1_1q<P64>(nfw_a1(n, k), n); }
G3L_HD double P64::u_a1(double a1, const Nfw &n) { return u_nfw_a1_1q<P64>(a1, n); }

// FP64 with the reference's own expressions for 10^x of l1, l2 and the r_200
// cube root (pow), the HOD occupations and NsatNsat's power
// (src/HOD.cu:35-93: log(m/M_th) and pow(m/M', beta) of the mass itself;
// Nsat's repeated erf term evaluated once).  The NNMap 3-halo term uses it:
// its squeezed bispectrum configurations (k3 -> 0, where the cosine-rule
// angles lose most of their digits) make that term's integrals sensitive to
// last-bit differences in any factor, and these forms, together with the
// projection factor in the reference's order, bring its values closest to
// the reference's (about half as many integrals move by more than 1e-4).
// The other terms refine identically either way and keep P64's cheaper
// forms.
struct P64R : P64 {
  // 10^x of l1, l2 and x^(1/3) of r_200 as the reference's pow
  // (src/NNMap_Model.cu:986-991, src/generalFunctions.cu:36)
  G3L_HD static double l(double x) { return pow(10.0, x); }
  G3L_HD static double cbrt3(double x) { return pow(x, 1. / 3.); }
  G3L_HD static Occ occ(const DevHod &h, double lth, double lmp, const Mass &M) {
    (void)lth;
    (void)lmp;
    double e = 1 + erf(log(M.m / h.mth) / h.sigma / 1.414213562);
    Occ o;
    o.nc = 0.5 * h.alpha * e;
    o.ns = 0.5 * e * pow(M.m / h.mprime, h.beta);
    return o;
  }
  G3L_HD static double mpow(const Mass &M, double eps, double lmmin) {
    (void)lmmin;
    return pow(M.m, eps);
  }
  // the library's log in Ci(x)
  G3L_HD static double ln_ci(double x) { return log(x); }
  G3L_HD static double u(const Nfw &n, double k) { return u_nfw<P64R>(n, k); }
  G3L_HD static double u_a1(double a1, const Nfw &n) { return u_nfw_a1<P64R>(a1, n); }
};

// u_NFW for a2 < 4 (small k r_s), the FP64 branch of P64::u on its own:
//   u = 1 + [(cos a1 - 1) ln(1+c) + (a1 c - sin a1 c)/a2
//            + cos a1 (Cr(a2) - Cr(a1)) + sin a1 (Si a2 - Si a1)] / F,
// over one common denominator as in u_nfw_a1_1q.  The Si/Ci differences of
// nearby arguments and the 1 - O(a1^2) structure need FP64 here; equal to
// P64::u bit for bit in this branch.
G3L_HD double u_small64(const Nfw &n, double a1, double a2) {
  double sx, cx, s1, c1, sn1, sd1, cn1, cd1, sn2, sd2, cn2, cd2;
  sincos_mm(a1 * n.c, sx, cx);
  sincos_mm(a1, s1, c1);
  sici_small_nd(a1, sn1, sd1, cn1, cd1);
  sici_small_nd(a2, sn2, sd2, cn2, cd2);
  const double C = cd1 * cd2, S = sd1 * sd2, CS = C * S;
  const double dC = fma(cn2, cd1, -cn1 * cd2);
  const double dS = fma(sn2, sd1, -sn1 * sd2);
  const double t0 = fma((c1 - 1) * n.lc, a2, a1 * n.c - sx);
  const double num = fma(t0, CS, a2 * fma(c1 * dC, S, s1 * dS * C));
  return 1 + fdiv(num, a2 * CS * n.F);
}

// ======================================================================
// Mixed policy: FP32 transcendental bulk.
namespace mix {

// 2^a for an FP64 argument: the integer part exactly, the fraction by exp2f.
G3L_HD double exp2_split(double a) {
  double n = floor(a);
  float f = (float)(a - n);
  return ldexp((double)exp2f(f), (int)n);
}
constexpr double kLog2e = 1.4426950408889634074;
G3L_HD double exp_split(double a) { return exp2_split(a * kLog2e); }

// sin and cos of an FP64 argument: reduction by 2 pi in FP64 (two-term
// Cody-Waite), FP32 evaluation of the remainder.
G3L_HD void sincos_red(double a, float &s, float &c) {
  const double inv2pi = 0.15915494309189533577;
  const double p1 = 6.28318530717958623200, p2 = 2.44929359829470635445e-16;
  double n = rint(a * inv2pi);
  double r = fma(-n, p1, a);
  r = fma(-n, p2, r);
  sincosf((float)r, &s, &c);
}

// Si/Ci auxiliary functions in FP32 for x >= 4, and f(x) - 1/x from the
// coefficient differences of its numerator and denominator (no cancellation).
G3L_HD void aux32(float x, float &f, float &g, float &fm) {
  float t = 1.0f / (x * x);
  float fn = fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t,
                 -4.94701168645415959931e11f, 4.94816688199951963482e12f),
                 1.00795182980368574617e13f), 4.20968180571076940208e12f),
                 6.40533830574022022911e11f), 4.33736238870432522765e10f),
                 1.43073403821274636888e9f), 2.37750310125431834034e7f),
                 1.96396372895146869801e5f), 7.44437068161936700618e2f);
  float fd = fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t,
                 1.11535493509914254097e13f, 1.43468549171581016479e13f),
                 5.06084464593475076774e12f), 7.08501308149515401563e11f),
                 4.58595115847765779830e10f), 1.47478952192985464958e9f),
                 2.41535670165126845144e7f), 1.97865247031583951450e5f),
                 7.46437068161927678031e2f);
  float gn = fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t,
                 -1.36517137670871689e12f, 6.43291613143049485e12f),
                 1.81004487464664575e13f), 7.57664583257834349e12f),
                 1.09049528450362786e12f), 6.83052205423625007e10f),
                 2.06297595146763354e9f), 3.12557570795778731e7f),
                 2.35239181626478200e5f), 8.1359520115168615e2f);
  float gd = fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t,
                 3.99653257887490811e13f, 4.01839087307656620e13f),
                 1.17164723371736605e13f), 1.39866710696414565e12f),
                 7.87465017341829930e10f), 2.23355543278099360e9f),
                 3.26026661647090822e7f), 2.40036752835578777e5f),
                 8.19595201151451564e2f);
  float dn = fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t, fmaf(t,
                 -4.94701168645415954590e11f, -6.20538246899190625000e12f),
                 -4.26733661912124414062e12f), -8.51162840223981323242e11f),
                 -6.79674775754933776855e10f), -2.48588769773332548141e9f),
                 -4.40554837171082794666e7f), -3.78536003969501121901e5f),
                 -1.46887413643708168820e3f), -1.99999999999097743952e0f);
  float den = 1.0f + t * fd;
  f = (1.0f + t * fn) / (den * x);
  g = (1.0f + t * gn) * t / (1.0f + t * gd);
  fm = t * dn / (den * x);
}

// Si(x) = x A(x^2)/B(x^2) and Cr(x) = Ci(x) - gamma - ln x = x^2 C(x^2)/D(x^2), x < 4.
G3L_HD void small32(float x, float &si, float &cr) {
  float x2 = x * x;
  float A = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2,
                -6.05338212010422477e-16f, 7.08240282274875911e-13f),
                -3.53201978997168357e-10f), 9.43280809438713025e-8f),
                -1.41018536821330254e-5f), 1.15457225751016682e-3f),
                -4.54393409816329991e-2f), 1.0f);
  float B = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2,
                3.21107051193712168e-16f, 4.5049097575386581e-13f),
                3.28067571055789734e-10f), 1.55654986308745614e-7f),
                4.99175116169755106e-5f), 1.01162145739225565e-2f), 1.0f);
  float C = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2,
                -9.93728488857585407e-15f, 1.06480802891189243e-11f),
                -4.68889508144848019e-9f), 1.05297363846239184e-6f),
                -1.27528342240267686e-4f), 7.51851524438898291e-3f), -0.25f);
  float D = fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2,
                1.39759616731376855e-18f, 1.89106054713059759e-15f),
                1.38536352772778619e-12f), 6.97071295760958946e-10f),
                2.55533277086129636e-7f), 6.72126800814254432e-5f),
                1.1592605689110735e-2f), 1.0f);
  si = x * A / B;
  cr = x2 * C / D;
}

}  // namespace mix

struct PMix {
  static constexpr bool kMixed = true;
  G3L_HD static double l(double x) { return exp10(x); }
  G3L_HD static double cbrt3(double x) { return cbrt(x); }

  // The HOD occupations and NsatNsat's power stay FP64 (log-space mass
  // exponentials and erf, as P64): they are evaluated once per (mass,
  // redshift) class of a region, so FP32 would save little, and their FP32
  // rounding moved whole integrals of this precision form by up to 2e-3.
  G3L_HD static Occ occ(const DevHod &h, double lth, double lmp, const Mass &M) {
    return P64::occ(h, lth, lmp, M);
  }
  G3L_HD static double mpow(const Mass &M, double eps, double lmmin) {
    return P64::mpow(M, eps, lmmin);
  }

  // Truncated-NFW transform in FP32 with the cancellations removed
  // algebraically (equal to the reference's formula in exact arithmetic):
  //   a1 >= 4:       u F = g(a1) + sin(a1 c) (f(a2) - 1/a2) - g(a2) cos(a1 c)
  //   a2 <  4:       u = 1 + [(cos a1 - 1) ln(1+c) + (a1 c - sin a1 c)/a2
  //                        + cos a1 (Cr(a2) - Cr(a1)) + sin a1 (Si(a2) - Si(a1))] / F
  //   a1 < 4 <= a2:  u F = pi/2 sin a1 - sin a1 Si(a1) - cos a1 Ci(a1)
  //                        + sin(a1 c) (f(a2) - 1/a2) - g(a2) cos(a1 c),
  //                  with Ci = gamma + ln a1 + Cr; its O(1) terms sum to
  //                  u F ~ 0.2-2 there, so FP32 keeps ~6 digits.
  G3L_HD static double u(const Nfw &n, double k) { return u_a1(nfw_a1(n, k), n); }
  G3L_HD static double u_a1(double a1d, const Nfw &n) {
    double a2d = a1d * (1 + n.c);
    // small k r_s: u = 1 - O(a1^2) from differences of nearby Si/Ci values,
    // kept in FP64
    if (a2d < 4.0) return u_small64(n, a1d, a2d);
    double xd = a1d * n.c;
    float sx, cxx;
    mix::sincos_red(xd, sx, cxx);
    if (a1d >= 4.0) {
      float f1, g1, fm1, f2, g2, fm2;
      mix::aux32((float)a1d, f1, g1, fm1);
      mix::aux32((float)a2d, f2, g2, fm2);
      float uF = g1 + sx * fm2 - g2 * cxx;
      return (double)(uF / (float)n.F);
    }
    float s1, c1;
    mix::sincos_red(a1d, s1, c1);
    float a1 = (float)a1d, a2 = (float)a2d;
    float si1, cr1;
    mix::small32(a1, si1, cr1);
    float f2, g2, fm2;
    mix::aux32(a2, f2, g2, fm2);
    const float ci1 = 0.5772156649f + logf(a1) + cr1;
    const float uF = 1.5707963267948966f * s1 - s1 * si1 - c1 * ci1 + sx * fm2 - g2 * cxx;
    return (double)(uF / (float)n.F);
  }

  G3L_HD static double ap(double eta) {
    double e2 = 0.5 * eta * eta;
    return e2 * mix::exp_split(-e2);
  }
};

// ======================================================================
// Shared structure of the integrands.

template <class P>
G3L_HD Occ occ1(const TermArgs &a, const Mass &M) {
  return P::occ(a.h1, a.lth1, a.lmp1, M);
}
template <class P>
G3L_HD Occ occ2(const TermArgs &a, const Mass &M) {
  return P::occ(a.h2, a.lth2, a.lmp2, M);
}

// src/HOD.cu:78-94 with scale1 = scale2 = 1.
template <class P>
G3L_HD double nsatnsat(const TermArgs &a, double ns1, double ns2,
                                           const Mass &M) {
  if (a.same) return (ns1 + 1.0 * 1.0 - 1) * ns1;
  double r = ns1 * ns2 + a.A * P::mpow(M, a.eps, a.lmmin) * sqrt(ns1 * ns2);
  return r < 0 ? 0.0 : r;
}

// F kernel with the reference's k2/k2 (src/generalFunctions.cu:89-92).
G3L_HD double fkern(double k1, double k2, double c) {
  return 0.7143 + 0.2857 * c * c + 0.5 * c * (k1 / k2 + k2 / k2);
}
// Tree-level bispectrum (src/generalFunctions.cu:94-147).
G3L_HD double bilin(const DevTables &T, double k1, double k2, double cosphi,
                                        double z) {
  double k3 = sqrt(k1 * k1 + k2 * k2 + 2 * k1 * k2 * cosphi);
  double c13 = (k2 * k2 - k3 * k3 - k1 * k1) / (2 * k1 * k3);
  double c23 = (k1 * k1 - k2 * k2 - k3 * k3) / (2 * k3 * k2);
  int zi = zix_round(T, z);
  double P1 = pk(T, zi, k1), P2 = pk(T, zi, k2), P3 = pk(T, zi, k3);
  return 2 * (fkern(k1, k2, cosphi) * P1 * P2 + fkern(k1, k3, c13) * P1 * P3 +
              fkern(k2, k3, c23) * P2 * P3);
}

// Coordinates shared by all terms: l1, l2, l3, cos phi, z and k = l/w.
struct Geo {
  double l1, l2, l3, cphi, z, w, k1, k2, k3;
  int zr, zt;
};
template <class P>
G3L_HD Geo make_geo(const DevTables &T, const double *x, int D) {
  Geo g;
  g.l1 = P::l(x[0]);
  g.l2 = P::l(x[1]);
  g.cphi = cos(x[2]);
  g.z = x[D - 1];
  g.zr = zix_round(T, g.z);
  g.zt = zix_trunc(T, g.z);
  g.w = T.w[g.zr];
  g.l3 = sqrt(g.l1 * g.l1 + g.l2 * g.l2 + 2 * g.l1 * g.l2 * g.cphi);
  g.k1 = g.l1 / g.w;
  g.k2 = g.l2 / g.w;
  g.k3 = g.l3 / g.w;
  return g;
}

// ======================================================================
// Per-kind structure: which coordinates are halo masses, and the u_NFW
// calls that depend on only part of the coordinates.  Coordinates:
//   5-D (1h): l1, l2, phi, m, z;  6-D (2h): l1, l2, phi, m1, m2, z;
//   7-D (3h): l1, l2, phi, m1, m2, m3, z  (l and m as log10).
// USpec {mask, kidx, slot, fsel}: u_NFW(k_kidx, mass slot, f) with f = 1,
// the first or the second HOD's f for fsel = 0, 1, 2; mask = coordinates it
// depends on (l_kidx, the mass, z; l1, l2, phi for k3).
struct USpec {
  unsigned mask;
  int kidx, slot, fsel;
};
constexpr unsigned kL1 = bit(0), kL2 = bit(1), kL3 = bit(0) | bit(1) | bit(2);

template <int KIND>
struct Spec;
template <>
struct Spec<kNN1> {
  static constexpr int D = 5, NM = 1, NU = 3, nss_slot = 0;
  static constexpr bool two_hods = true, has_bilin = false, has_pk = false;
  static constexpr int mdim[3] = {3, 0, 0};
  static constexpr USpec u[6] = {{kL1 | bit(3) | bit(4), 1, 0, 1},   // u(k1, m, f1)
                                 {kL2 | bit(3) | bit(4), 2, 0, 2},   // u(k2, m, f2)
                                 {kL3 | bit(3) | bit(4), 3, 0, 0}};  // u(k3, m), every point
};
template <>
struct Spec<kNN2> {
  static constexpr int D = 6, NM = 2, NU = 6, nss_slot = 1;
  static constexpr bool two_hods = true, has_bilin = false, has_pk = true;
  static constexpr int mdim[3] = {3, 4, 0};
  static constexpr USpec u[6] = {{kL1 | bit(4) | bit(5), 1, 1, 1},   // u(k1, m2, f1)
                                 {kL2 | bit(4) | bit(5), 2, 1, 2},   // u(k2, m2, f2)
                                 {kL1 | bit(3) | bit(5), 1, 0, 1},   // u(k1, m1, f1)
                                 {kL2 | bit(3) | bit(5), 2, 0, 2},   // u(k2, m1, f2)
                                 {kL3 | bit(3) | bit(5), 3, 0, 0},   // u(k3, m1)
                                 {kL3 | bit(4) | bit(5), 3, 1, 0}};  // u(k3, m2)
};
template <>
struct Spec<kNN3> {
  static constexpr int D = 7, NM = 3, NU = 3, nss_slot = -1;
  static constexpr bool two_hods = true, has_bilin = true, has_pk = false;
  static constexpr int mdim[3] = {3, 4, 5};
  static constexpr USpec u[6] = {{kL3 | bit(5) | bit(6), 3, 2, 0},   // u(k3, m3)
                                 {kL1 | bit(3) | bit(6), 1, 0, 1},   // u(k1, m1, f1)
                                 {kL2 | bit(4) | bit(6), 2, 1, 2}};  // u(k2, m2, f2)
};
template <>
struct Spec<kNM1> {
  static constexpr int D = 5, NM = 1, NU = 3, nss_slot = -1;
  static constexpr bool two_hods = false, has_bilin = false, has_pk = false;
  static constexpr int mdim[3] = {3, 0, 0};
  static constexpr USpec u[6] = {{kL1 | bit(3) | bit(4), 1, 0, 1},   // u(k1, m, f)
                                 {kL2 | bit(3) | bit(4), 2, 0, 0},   // u(k2, m)
                                 {kL3 | bit(3) | bit(4), 3, 0, 0}};  // u(k3, m), every point
};
template <>
struct Spec<kNM2> {
  static constexpr int D = 6, NM = 2, NU = 6, nss_slot = -1;
  static constexpr bool two_hods = false, has_bilin = false, has_pk = true;
  static constexpr int mdim[3] = {3, 4, 0};
  static constexpr USpec u[6] = {{kL2 | bit(3) | bit(5), 2, 0, 0},   // u(k2, m1)
                                 {kL3 | bit(3) | bit(5), 3, 0, 0},   // u(k3, m1)
                                 {kL2 | bit(4) | bit(5), 2, 1, 0},   // u(k2, m2)
                                 {kL3 | bit(4) | bit(5), 3, 1, 0},   // u(k3, m2)
                                 {kL1 | bit(4) | bit(5), 1, 1, 1},   // u(k1, m2, f)
                                 {kL1 | bit(3) | bit(5), 1, 0, 1}};  // u(k1, m1, f)
};
template <>
struct Spec<kNM3> {
  static constexpr int D = 7, NM = 3, NU = 3, nss_slot = -1;
  static constexpr bool two_hods = false, has_bilin = true, has_pk = false;
  static constexpr int mdim[3] = {3, 4, 5};
  static constexpr USpec u[6] = {{kL2 | bit(3) | bit(6), 2, 0, 0},   // u(k2, m1)
                                 {kL3 | bit(4) | bit(6), 3, 1, 0},   // u(k3, m2)
                                 {kL1 | bit(4) | bit(6), 1, 1, 1}};  // u(k1, m2, f)
};
// Whether the integrand of KIND needs u_NFW's normalisation (c, ln(1+c), F)
// for mass slot `slot` with f selector `fsel`: one of its u_NFW factors uses
// it.
template <int KIND>
constexpr bool uses_nfw(int slot, int fsel) {
  for (int i = 0; i < Spec<KIND>::NU; ++i)
    if (Spec<KIND>::u[i].slot == slot && Spec<KIND>::u[i].fsel == fsel) return true;
  return false;
}

// Coordinates of the tree-level bispectrum: l1, l2, phi, z.
template <int D>
constexpr unsigned bilin_mask() { return kL3 | bit(D - 1); }

G3L_HD double f_of(const TermArgs &a, int fsel) {
  return fsel == 0 ? 1.0 : fsel == 1 ? a.h1.f : a.h2.f;
}

// Redshift-only factors of the integrands, per redshift index zi:
//   zpre = 1/n1/n2/rho (NNMap), 1/n/rho/rho (NMapMap), the leading factor of B;
//   zfac = g p1 p2 (1+z) / dwdz / w^3 (NNMap), g^2 p1 (1+z)^2 / dwdz / w^2
//   (NMapMap), the projection factor b = zfac B (the NNMap 3-halo term
//   applies it in the reference's order instead, proj_3h).
template <int KIND>
G3L_HD void zfactors(const DevTables &T, const TermArgs &a, int zi, double z, double w,
                     double &zpre, double &zfac) {
  if constexpr (KIND <= kNN3) {
    zpre = 1. / a.nb1[zi] / a.nb2[zi] / T.rho[zi];
    zfac = 1. / T.dwdz[zi] * T.g[zi] * T.p1[zi] * T.p2[zi] / w / w / w * (1. + z);
  } else {
    const double rho = T.rho[zi];
    zpre = 1. / a.nb1[zi] / rho / rho;
    zfac = 1. / T.dwdz[zi] * T.g[zi] * T.g[zi] * T.p1[zi] / w / w * (1. + z) * (1. + z);
  }
}

// The evaluation policy of an integrand: P64R for the NNMap 3-halo term and
// the n_bar integrand in FP64, the requested policy otherwise.
template <int KIND, class P>
struct EvalPolicy {
  using type = P;
};
template <>
struct EvalPolicy<kNN3, P64> {
  using type = P64R;
};
// n_bar(z), a factor of every term, with the reference's HOD expressions too
template <>
struct EvalPolicy<kNbar, P64> {
  using type = P64R;
};

// a / d correctly rounded from rd = RN(1/d): q = RN(a rd), then one FMA
// residual correction (Markstein's theorem: with rd the correctly rounded
// reciprocal the corrected quotient is RN(a / d) for normal operands, no
// overflow or underflow).  Three instructions instead of the IEEE division
// sequence, for divisors that repeat over many points.
G3L_HD double div_rn(double a, double d, double rd) {
  const double q = a * rd;
  return fma(fma(-d, q, a), rd, q);
}

// The NNMap 3-halo projection b = B / dwdz g p1 p2 / w^3 (1+z) in the
// reference's order (src/NNMap_Model.cu:1031-1033), from the divisors and
// their correctly rounded reciprocals; equal to the IEEE-division form.  The
// other terms use the factor zfac of zfactors: b = zfac B.
G3L_HD double proj_3h(double B, double dwdz, double rdwdz, double g, double p1, double p2,
                      double w, double rw, double z) {
  double b = div_rn(B, dwdz, rdwdz) * g * p1 * p2;
  b = div_rn(b, w, rw);
  b = div_rn(b, w, rw);
  b = div_rn(b, w, rw);
  return b * (1. + z);
}

// Direct source: every quantity computed at the point itself.
template <int KIND, class P>
struct DirectSrc {
  using S = Spec<KIND>;
  const DevTables &T;
  const TermArgs &a;
  Geo G;
  Mass M[S::NM];
  G3L_HD DirectSrc(const DevTables &T_, const TermArgs &a_, const double *x) : T(T_), a(a_) {
    G = make_geo<P>(T, x, S::D);
#pragma unroll
    for (int i = 0; i < S::NM; ++i) M[i] = make_mass<P>(T, x[S::mdim[i]], G.zt);
  }
  G3L_HD const Geo &geo() const { return G; }
  G3L_HD Mass mass(int i) const { return M[i]; }
  G3L_HD Occ occ1(int i) const { return g3l::occ1<P>(a, M[i]); }
  G3L_HD Occ occ2(int i) const { return g3l::occ2<P>(a, M[i]); }
  G3L_HD double nss(int i) const { return nsatnsat<P>(a, occ1(i).ns, occ2(i).ns, M[i]); }
  G3L_HD Nfw nfw(int slot, int fsel) const { return make_nfw(T, M[slot], G.zt, f_of(a, fsel)); }
  template <int I>
  G3L_HD double u() const {
    constexpr USpec s = S::u[I];
    const double k = s.kidx == 1 ? G.k1 : s.kidx == 2 ? G.k2 : G.k3;
    return P::u(nfw(s.slot, s.fsel), k);
  }
  G3L_HD double ap1() const { return P::ap(a.t1 * G.l1); }
  G3L_HD double ap2() const { return P::ap(a.t2 * G.l2); }
  G3L_HD double bl() const { return bilin(T, G.k1, G.k2, G.cphi, G.z); }
  G3L_HD double ap3() const { return P::ap(a.t3 * G.l3); }
  G3L_HD double pk3() const { return pk(T, G.zr, G.k3); }
  G3L_HD double pk1() const { return pk(T, G.zr, G.k1); }
  G3L_HD double pk2() const { return pk(T, G.zr, G.k2); }
  G3L_HD double zpre() const {
    double p, f;
    zfactors<KIND>(T, a, G.zr, G.z, G.w, p, f);
    return p;
  }
  G3L_HD double proj(double B) const {
    const int zi = G.zr;
    if constexpr (KIND == kNN3) {
      return B / T.dwdz[zi] * T.g[zi] * T.p1[zi] * T.p2[zi] / G.w / G.w / G.w * (1. + G.z);
    } else {
      double p, f;
      zfactors<KIND>(T, a, zi, G.z, G.w, p, f);
      return f * B;
    }
  }
};

// The six integrands over a value source S (DirectSrc, or the kernel's
// per-region cache); the arithmetic after the look-ups is the same for both.

// NNMap 1-halo (src/NNMap_Model.cu:544-593).
template <class P, class S>
G3L_HD double nnmap_1h(const DevTables &T, const TermArgs &a, const S &s) {
  const Geo &G = s.geo();
  const Mass M = s.mass(0);
  const int zi = G.zr;
  Occ o1 = s.occ1(0), o2 = s.occ2(0);
  double nss = s.nss(0);
  double u3 = s.template u<2>();
  double u1 = s.template u<0>();
  double u2 = s.template u<1>();
  double m = M.m, l1 = G.l1, l2 = G.l2;
  double B = s.zpre() * T.hmf[zi * T.nb + M.ir] * m * u3;
  B *= o1.nc * o2.ns * u2 + o2.nc * o1.ns * u1 + nss * u1 * u2;
  double b = s.proj(B);
  return l1 * l2 * m * l1 * l2 * s.ap1() * s.ap2() * s.ap3() * b;
}

// NNMap 2-halo (src/NNMap_Model.cu:787-856).
template <class P, class S>
G3L_HD double nnmap_2h(const DevTables &T, const TermArgs &a, const S &s) {
  const Geo &G = s.geo();
  const Mass M1 = s.mass(0), M2 = s.mass(1);
  const int zi = G.zr, nb = T.nb;
  double P1 = s.pk1(), P2 = s.pk2(), P3 = s.pk3();
  // G_gg(k1, k2, m2)
  Occ a1 = s.occ1(1), a2 = s.occ2(1);
  double nss = s.nss(1);
  double u_k1_m2_f1 = s.template u<0>();
  double u_k2_m2_f2 = s.template u<1>();
  double u_k1_m1_f1 = s.template u<2>();
  double u_k2_m1_f2 = s.template u<3>();
  double Ggg = a1.nc * a2.ns * u_k2_m2_f2 + a2.nc * a1.ns * u_k1_m2_f1 +
               nss * u_k1_m2_f1 * u_k2_m2_f2;
  Occ b1 = s.occ1(0), b2 = s.occ2(0);
  double G_k1_m1_h1 = b1.nc + b1.ns * u_k1_m1_f1;
  double G_k2_m2_h2 = a2.nc + a2.ns * u_k2_m2_f2;
  double G_k2_m1_h2 = b2.nc + b2.ns * u_k2_m1_f2;
  double G_k1_m2_h1 = a1.nc + a1.ns * u_k1_m2_f1;
  double u_k3_m1 = s.template u<4>();
  double u_k3_m2 = s.template u<5>();
  double m1 = M1.m, m2 = M2.m, l1 = G.l1, l2 = G.l2;
  double B = s.zpre() * T.hmf[zi * nb + M1.ir] *
             T.hmf[zi * nb + M2.ir] * T.bh[zi * nb + M1.ir] * T.bh[zi * nb + M2.ir] *
             (m1 * u_k3_m1 * Ggg * P3 + m2 * u_k3_m2 * G_k1_m1_h1 * G_k2_m2_h2 * P1 +
              m2 * u_k3_m2 * G_k2_m1_h2 * G_k1_m2_h1 * P2);
  double b = s.proj(B);
  return l1 * l2 * m1 * m2 * l1 * l2 * s.ap1() * s.ap2() * s.ap3() * b;
}

// NNMap 3-halo (src/NNMap_Model.cu:1003-1037).
template <class P, class S>
G3L_HD double nnmap_3h(const DevTables &T, const TermArgs &a, const S &s) {
  const Geo &G = s.geo();
  const Mass M1 = s.mass(0), M2 = s.mass(1), M3 = s.mass(2);
  const int zi = G.zr, nb = T.nb;
  double u3 = s.template u<0>();
  Occ o1 = s.occ1(0);
  double G1 = o1.nc + o1.ns * s.template u<1>();
  Occ o2 = s.occ2(1);
  double G2 = o2.nc + o2.ns * s.template u<2>();
  double m1 = M1.m, m2 = M2.m, m3 = M3.m, l1 = G.l1, l2 = G.l2;
  double B = s.zpre() * T.hmf[zi * nb + M1.ir] *
             T.hmf[zi * nb + M2.ir] * T.hmf[zi * nb + M3.ir] * m3 * u3 * G1 * G2 * s.bl() *
             T.bh[zi * nb + M1.ir] * T.bh[zi * nb + M2.ir] * T.bh[zi * nb + M3.ir];
  double b = s.proj(B);
  return l1 * l2 * m1 * m2 * m3 * l1 * l2 * s.ap1() * s.ap2() * s.ap3() * b;
}

// NMapMap 1-halo (src/NMapMap_Model.cu:263-301).
template <class P, class S>
G3L_HD double nmapmap_1h(const DevTables &T, const TermArgs &a, const S &s) {
  const Geo &G = s.geo();
  const Mass M = s.mass(0);
  const int zi = G.zr;
  double nbar = a.nb1[zi];
  Occ o = s.occ1(0);
  double uf = s.template u<0>();
  double m = M.m, l1 = G.l1, l2 = G.l2;
  double B = s.zpre() * T.hmf[zi * T.nb + M.ir] * m * m * (o.nc + o.ns * uf) *
             s.template u<1>() * s.template u<2>();
  double b = s.proj(B);
  return l1 * l2 * m * l1 * l2 * s.ap1() * s.ap2() * s.ap3() * b / nbar;
}

// NMapMap 2-halo (src/NMapMap_Model.cu:402-468).
template <class P, class S>
G3L_HD double nmapmap_2h(const DevTables &T, const TermArgs &a, const S &s) {
  const Geo &G = s.geo();
  const Mass M1 = s.mass(0), M2 = s.mass(1);
  const int zi = G.zr, nb = T.nb;
  double nbar = a.nb1[zi];
  double P1 = s.pk1(), P2 = s.pk2(), P3 = s.pk3();
  Occ o1 = s.occ1(0), o2 = s.occ1(1);
  double u_k2_m1 = s.template u<0>(), u_k3_m1 = s.template u<1>();
  double u_k2_m2 = s.template u<2>(), u_k3_m2 = s.template u<3>();
  double u_k1_m2_f = s.template u<4>();
  double u_k1_m1_f = s.template u<5>();
  double m1 = M1.m, m2 = M2.m, l1 = G.l1, l2 = G.l2;
  double B = s.zpre() * T.hmf[zi * nb + M1.ir] * T.hmf[zi * nb + M2.ir] *
             T.bh[zi * nb + M1.ir] * T.bh[zi * nb + M2.ir] *
             (m1 * m2 * u_k2_m1 * u_k3_m2 * (o2.nc + o2.ns * u_k1_m2_f) * P2 +
              m1 * m2 * u_k3_m1 * u_k2_m2 * (o2.nc + o2.ns * u_k1_m2_f) * P3 +
              m2 * m2 * u_k2_m1 * u_k3_m2 * (o1.nc + o1.ns * u_k1_m1_f) * P1);
  double b = s.proj(B);
  return l1 * l2 * m1 * m2 * l1 * l2 * s.ap1() * s.ap2() * s.ap3() * b / nbar;
}

// NMapMap 3-halo: Ncen, Nsat at m3 with u_NFW(k1, m2, f) (src/NMapMap_Model.cu:593-632).
template <class P, class S>
G3L_HD double nmapmap_3h(const DevTables &T, const TermArgs &a, const S &s) {
  const Geo &G = s.geo();
  const Mass M1 = s.mass(0), M2 = s.mass(1), M3 = s.mass(2);
  const int zi = G.zr, nb = T.nb;
  double nbar = a.nb1[zi];
  Occ o = s.occ1(2);
  double u_k2_m1 = s.template u<0>();
  double u_k3_m2 = s.template u<1>();
  double u_k1_m2_f = s.template u<2>();
  double m1 = M1.m, m2 = M2.m, m3 = M3.m;
  double l1 = G.l1, l2 = G.l2;
  double B = s.zpre() * T.hmf[zi * nb + M1.ir] * T.hmf[zi * nb + M2.ir] *
             T.hmf[zi * nb + M3.ir] * T.bh[zi * nb + M1.ir] * T.bh[zi * nb + M2.ir] *
             T.bh[zi * nb + M3.ir] * m1 * m2 * u_k2_m1 * u_k3_m2 * (o.nc + o.ns * u_k1_m2_f) *
             s.bl();
  double b = s.proj(B);
  return l1 * l2 * m1 * m2 * m3 * l1 * l2 * s.ap1() * s.ap2() * s.ap3() * b / nbar;
}

// The integrand of kind KIND at one point (direct evaluation).
template <int KIND, class P, class S>
G3L_HD double integrand_src(const DevTables &T, const TermArgs &a, const S &s) {
  if constexpr (KIND == kNN1) {
    return nnmap_1h<P>(T, a, s);
  } else if constexpr (KIND == kNN2) {
    return nnmap_2h<P>(T, a, s);
  } else if constexpr (KIND == kNN3) {
    return nnmap_3h<P>(T, a, s);
  } else if constexpr (KIND == kNM1) {
    return nmapmap_1h<P>(T, a, s);
  } else if constexpr (KIND == kNM2) {
    return nmapmap_2h<P>(T, a, s);
  } else {
    return nmapmap_3h<P>(T, a, s);
  }
}

// n_bar integrand, x = log10 m, theta slot 1 = z (src/generalFunctions.cu:233-244);
// the table-precomputation integrands share its 1-D kind (mode != kOneDNbar).
template <class P>
G3L_HD double nbar_kernel(const DevTables &T, const TermArgs &a,
                                              const double *x) {
  if (a.mode != kOneDNbar) return precomp_kernel(a, x[0]);
  double z = a.t1;
  int zr = zix_round(T, z);
  Mass M = make_mass<P>(T, x[0], zix_trunc(T, z));
  Occ o = occ1<P>(a, M);
  return M.m * T.hmf[zr * T.nb + M.ir] * (o.nc + o.ns);
}

}  // namespace g3l
