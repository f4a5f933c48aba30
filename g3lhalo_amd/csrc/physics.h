// physics.h -- FP64 device integrands of the NNMap / NMapMap path (gfx950).
//
// Each function follows the reference's expression order so that, apart from
// the last-place differences between device and host libm, the values equal
// the reference CPU build's.  Identical sub-expressions that the reference
// evaluates more than once (u_NFW with equal arguments in G_gg/G_g, Ncen and
// Nsat sharing one erf, r_200 per mass) are evaluated once: recomputing a
// pure function yields the same bits, so the sharing changes nothing.
//
// Reference: src/generalFunctions.cu:6-147, src/HOD.cu:35-94,
//            src/NNMap_Model.cu:544-1108, src/NMapMap_Model.cu:263-660.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

#pragma clang fp contract(off)

namespace g3l {

// Tables resident in HBM (z-major 2-D tables, nb*nb + nb for P_lin).
struct DevTables {
  int nb;
  double zmin, zmax, kmin, kmax, mmin, mmax;
  double inv_dz;     // unused by the FP64 path (kept for the mixed path)
  const double *g, *p1, *p2, *w, *dwdz, *hmf, *plin, *bh, *conc, *rho;
};

struct DevHod {
  double f, alpha, mth, sigma, mprime, beta;
};

// ---------------------------------------------------------------- indices
// Nearest-index look-ups: round() in the integrands, int() truncation in
// u_NFW and r_200 (src/generalFunctions.cu:11-12,35).  Every index is clamped
// to the allocated table; inside the integration box the clamp never acts,
// it only keeps invalid input (e.g. a zero comoving distance) from reading
// outside the tables.
__device__ __forceinline__ int clampi(int i, int hi) { return i < 0 ? 0 : (i > hi ? hi : i); }
__device__ __forceinline__ int zix_round(const DevTables &T, double z) {
  return clampi((int)round((z - T.zmin) * T.nb / (T.zmax - T.zmin)), T.nb - 1);
}
__device__ __forceinline__ int zix_trunc(const DevTables &T, double z) {
  return clampi((int)((z - T.zmin) * T.nb / (T.zmax - T.zmin)), T.nb - 1);
}
__device__ __forceinline__ int mix_round(const DevTables &T, double m) {
  return clampi((int)round(log(m / T.mmin) * T.nb / log(T.mmax / T.mmin)), T.nb - 1);
}
__device__ __forceinline__ int mix_trunc(const DevTables &T, double m) {
  return clampi((int)(log(m / T.mmin) * T.nb / log(T.mmax / T.mmin)), T.nb - 1);
}
__device__ __forceinline__ int kix_round(const DevTables &T, double k) {
  return (int)round(log(k / T.kmin) * T.nb / log(T.kmax / T.kmin));
}
// P(k) look-up, linear below kmin, no upper guard (src/NNMap_Model.cu:819-839):
// an index past the row reads on into the next row; P_lin carries nb
// trailing zeros for the last row.
__device__ __forceinline__ double pk(const DevTables &T, int zi, double k) {
  int ki = kix_round(T, k);
  if (ki >= 0) return T.plin[clampi(zi * T.nb + ki, T.nb * T.nb + T.nb - 1)];
  return T.plin[zi * T.nb] / T.kmin * k;
}

// ----------------------------------------------------------------- Si, Ci
// Rational approximations of src/generalFunctions.cu:44-87.
__device__ __forceinline__ void sici(double x, double &si, double &ci) {
  double x2 = x * x;
  double x4 = x2 * x2;
  double x6 = x2 * x4;
  double x8 = x4 * x4;
  double x10 = x8 * x2;
  double x12 = x6 * x6;
  double x14 = x12 * x2;
  if (x < 4) {
    double a = 1 - 4.54393409816329991e-2 * x2 + 1.15457225751016682e-3 * x4 -
               1.41018536821330254e-5 * x6 + 9.43280809438713025e-8 * x8 -
               3.53201978997168357e-10 * x10 + 7.08240282274875911e-13 * x12 -
               6.05338212010422477e-16 * x14;
    double b = 1 + 1.01162145739225565e-2 * x2 + 4.99175116169755106e-5 * x4 +
               1.55654986308745614e-7 * x6 + 3.28067571055789734e-10 * x8 +
               4.5049097575386581e-13 * x10 + 3.21107051193712168e-16 * x12;
    si = x * a / b;
    a = -0.25 + 7.51851524438898291e-3 * x2 - 1.27528342240267686e-4 * x4 +
        1.05297363846239184e-6 * x6 - 4.68889508144848019e-9 * x8 +
        1.06480802891189243e-11 * x10 - 9.93728488857585407e-15 * x12;
    b = 1 + 1.1592605689110735e-2 * x2 + 6.72126800814254432e-5 * x4 +
        2.55533277086129636e-7 * x6 + 6.97071295760958946e-10 * x8 +
        1.38536352772778619e-12 * x10 + 1.89106054713059759e-15 * x12 +
        1.39759616731376855e-18 * x14;
    ci = 0.5772156649 + log(x) + x2 * a / b;
  } else {
    double x16 = x8 * x8;
    double x18 = x16 * x2;
    double x20 = x10 * x10;
    double cx, sx;
    sincos(x, &sx, &cx);
    double f = (1 + 7.44437068161936700618e2 / x2 + 1.96396372895146869801e5 / x4 +
                2.37750310125431834034e7 / x6 + 1.43073403821274636888e9 / x8 +
                4.33736238870432522765e10 / x10 + 6.40533830574022022911e11 / x12 +
                4.20968180571076940208e12 / x14 + 1.00795182980368574617e13 / x16 +
                4.94816688199951963482e12 / x18 - 4.94701168645415959931e11 / x20) /
               (1 + 7.46437068161927678031e2 / x2 + 1.97865247031583951450e5 / x4 +
                2.41535670165126845144e7 / x6 + 1.47478952192985464958e9 / x8 +
                4.58595115847765779830e10 / x10 + 7.08501308149515401563e11 / x12 +
                5.06084464593475076774e12 / x14 + 1.43468549171581016479e13 / x16 +
                1.11535493509914254097e13 / x18) /
               x;
    double g = (1 + 8.1359520115168615e2 / x2 + 2.35239181626478200e5 / x4 +
                3.12557570795778731e7 / x6 + 2.06297595146763354e9 / x8 +
                6.83052205423625007e10 / x10 + 1.09049528450362786e12 / x12 +
                7.57664583257834349e12 / x14 + 1.81004487464664575e13 / x16 +
                6.43291613143049485e12 / x18 - 1.36517137670871689e12 / x20) /
               (1 + 8.19595201151451564e2 / x2 + 2.40036752835578777e5 / x4 +
                3.26026661647090822e7 / x6 + 2.23355543278099360e9 / x8 +
                7.87465017341829930e10 / x10 + 1.39866710696414565e12 / x12 +
                1.17164723371736605e13 / x14 + 4.01839087307656620e13 / x16 +
                3.99653257887490811e13 / x18) /
               x2;
    si = 0.5 * 3.14159265358979323846 - f * cx - g * sx;
    ci = f * sx - g * cx;
  }
}

__device__ __forceinline__ double aperture(double eta) {
  return 0.5 * eta * eta * exp(-0.5 * eta * eta);
}

// ---------------------------------------------------------------- HOD
// Ncen and Nsat share log(m/M_th)/sigma/sqrt2 and its erf (src/HOD.cu:35-76).
struct Occ {
  double nc, ns;
};
__device__ __forceinline__ Occ occupation(const DevHod &h, double m) {
  double e = 1 + erf(log(m / h.mth) / h.sigma / 1.414213562);
  Occ o;
  o.nc = 0.5 * h.alpha * e;
  o.ns = 0.5 * e * pow(m / h.mprime, h.beta);
  return o;
}
// src/HOD.cu:78-94 with scale1 = scale2 = 1.
__device__ __forceinline__ double nsatnsat(double ns1, double ns2, bool same, double A,
                                           double eps, double m) {
  if (same) return (ns1 + 1.0 * 1.0 - 1) * ns1;
  double r = ns1 * ns2 + A * pow(m, eps) * sqrt(ns1 * ns2);
  return r < 0 ? 0.0 : r;
}

// --------------------------------------------------------------- u_NFW
// Per (m, z, f) part of the truncated-NFW transform (src/generalFunctions.cu:6-37).
struct Nfw {
  double r200, c, F;
};
__device__ __forceinline__ double r200(const DevTables &T, double m, double z) {
  int zi = zix_trunc(T, z);
  return pow(0.239 * m / T.rho[zi] / 200, 1. / 3.);
}
__device__ __forceinline__ Nfw nfw_setup(const DevTables &T, double m, double z, double f,
                                         double r) {
  int zi = zix_trunc(T, z);
  int mi = mix_trunc(T, m);
  Nfw n;
  n.r200 = r;
  n.c = f * T.conc[zi * T.nb + mi];
  n.F = log(1. + n.c) - n.c / (1. + n.c);
  return n;
}
__device__ __forceinline__ double unfw(const Nfw &n, double k) {
  double arg1 = k * n.r200 / n.c;
  double arg2 = arg1 * (1 + n.c);
  double si1, ci1, si2, ci2;
  sici(arg1, si1, ci1);
  sici(arg2, si2, ci2);
  double s1, c1;
  sincos(arg1, &s1, &c1);
  double term1 = s1 * (si2 - si1);
  double term2 = c1 * (ci2 - ci1);
  double term3 = -sin(arg1 * n.c) / arg2;
  return (term1 + term2 + term3) / n.F;
}

// ---------------------------------------------------------- Bi_lin, F
__device__ __forceinline__ double fkern(double k1, double k2, double c) {
  return 0.7143 + 0.2857 * c * c + 0.5 * c * (k1 / k2 + k2 / k2);
}
__device__ __forceinline__ double bilin(const DevTables &T, double k1, double k2, double cosphi,
                                        double z) {
  double k3 = sqrt(k1 * k1 + k2 * k2 + 2 * k1 * k2 * cosphi);
  double c13 = (k2 * k2 - k3 * k3 - k1 * k1) / (2 * k1 * k3);
  double c23 = (k1 * k1 - k2 * k2 - k3 * k3) / (2 * k3 * k2);
  int zi = zix_round(T, z);
  double P1 = pk(T, zi, k1), P2 = pk(T, zi, k2), P3 = pk(T, zi, k3);
  return 2 * (fkern(k1, k2, cosphi) * P1 * P2 + fkern(k1, k3, c13) * P1 * P3 +
              fkern(k2, k3, c23) * P2 * P3);
}

// ------------------------------------------------------------ term context
// Per-integral parameters seen by the integrand.
struct TermArgs {
  double t1, t2, t3;        // apertures [rad]
  DevHod h1, h2;            // HODs selected by type1, type2 (NMapMap: h1)
  const double *nb1, *nb2;  // n_bar selected by type (NMapMap: nb1)
  double A, eps;
  int same;                 // NsatNsat same-type branch
};

// NNMap 1-halo, x = (log10 l1, log10 l2, phi, log10 m, z) (src/NNMap_Model.cu:544-593).
__device__ __forceinline__ double nnmap_1h(const DevTables &T, const TermArgs &a,
                                           const double *x) {
  double l1 = pow(10.0, x[0]), l2 = pow(10.0, x[1]), phi = x[2], m = pow(10.0, x[3]);
  double z = x[4];
  int zi = zix_round(T, z);
  int mi = mix_round(T, m);
  double w_ = T.w[zi];
  double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi));
  double k1 = l1 / w_, k2 = l2 / w_, k3 = l3 / w_;
  double r = r200(T, m, z);
  Occ o1 = occupation(a.h1, m), o2 = occupation(a.h2, m);
  double nss = nsatnsat(o1.ns, o2.ns, a.same, a.A, a.eps, m);
  Nfw n1 = nfw_setup(T, m, z, 1, r);
  Nfw nf1 = nfw_setup(T, m, z, a.h1.f, r);
  Nfw nf2 = nfw_setup(T, m, z, a.h2.f, r);
  double u3 = unfw(n1, k3);
  double u1 = unfw(nf1, k1);
  double u2 = unfw(nf2, k2);
  double B = 1. / a.nb1[zi] / a.nb2[zi] / T.rho[zi] * T.hmf[zi * T.nb + mi] * m * u3;
  B *= o1.nc * o2.ns * u2 + o2.nc * o1.ns * u1 + nss * u1 * u2;
  double b = B / T.dwdz[zi] * T.g[zi] * T.p1[zi] * T.p2[zi] / w_ / w_ / w_ * (1. + z);
  return l1 * l2 * m * l1 * l2 * aperture(a.t1 * l1) * aperture(a.t2 * l2) *
         aperture(a.t3 * l3) * b;
}

// NNMap 2-halo, x = (log10 l1, log10 l2, phi, log10 m1, log10 m2, z)
// (src/NNMap_Model.cu:787-856).
__device__ __forceinline__ double nnmap_2h(const DevTables &T, const TermArgs &a,
                                           const double *x) {
  double l1 = pow(10.0, x[0]), l2 = pow(10.0, x[1]), phi = x[2];
  double m1 = pow(10.0, x[3]), m2 = pow(10.0, x[4]), z = x[5];
  const int nb = T.nb;
  int zi = zix_round(T, z);
  int i1 = mix_round(T, m1), i2 = mix_round(T, m2);
  double w_ = T.w[zi];
  double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi));
  double k1 = l1 / w_, k2 = l2 / w_, k3 = l3 / w_;
  double P1 = pk(T, zi, k1), P2 = pk(T, zi, k2), P3 = pk(T, zi, k3);
  double ra = r200(T, m1, z), rb = r200(T, m2, z);
  // G_gg(k1, k2, m2)
  Occ a1 = occupation(a.h1, m2), a2 = occupation(a.h2, m2);
  double nss = nsatnsat(a1.ns, a2.ns, a.same, a.A, a.eps, m2);
  Nfw m2f1 = nfw_setup(T, m2, z, a.h1.f, rb), m2f2 = nfw_setup(T, m2, z, a.h2.f, rb);
  Nfw m1f1 = nfw_setup(T, m1, z, a.h1.f, ra), m1f2 = nfw_setup(T, m1, z, a.h2.f, ra);
  double u_k1_m2_f1 = unfw(m2f1, k1), u_k2_m2_f2 = unfw(m2f2, k2);
  double u_k1_m1_f1 = unfw(m1f1, k1), u_k2_m1_f2 = unfw(m1f2, k2);
  double Ggg = a1.nc * a2.ns * u_k2_m2_f2 + a2.nc * a1.ns * u_k1_m2_f1 +
               nss * u_k1_m2_f1 * u_k2_m2_f2;
  Occ b1 = occupation(a.h1, m1), b2 = occupation(a.h2, m1);
  double G_k1_m1_h1 = b1.nc + b1.ns * u_k1_m1_f1;
  double G_k2_m2_h2 = a2.nc + a2.ns * u_k2_m2_f2;
  double G_k2_m1_h2 = b2.nc + b2.ns * u_k2_m1_f2;
  double G_k1_m2_h1 = a1.nc + a1.ns * u_k1_m2_f1;
  double u_k3_m1 = unfw(nfw_setup(T, m1, z, 1, ra), k3);
  double u_k3_m2 = unfw(nfw_setup(T, m2, z, 1, rb), k3);
  double B = 1. / a.nb1[zi] / a.nb2[zi] / T.rho[zi] * T.hmf[zi * nb + i1] *
             T.hmf[zi * nb + i2] * T.bh[zi * nb + i1] * T.bh[zi * nb + i2] *
             (m1 * u_k3_m1 * Ggg * P3 + m2 * u_k3_m2 * G_k1_m1_h1 * G_k2_m2_h2 * P1 +
              m2 * u_k3_m2 * G_k2_m1_h2 * G_k1_m2_h1 * P2);
  double b = 1. / T.dwdz[zi] * T.g[zi] * T.p1[zi] * T.p2[zi] / w_ / w_ / w_ * (1. + z) * B;
  return l1 * l2 * m1 * m2 * l1 * l2 * aperture(a.t1 * l1) * aperture(a.t2 * l2) *
         aperture(a.t3 * l3) * b;
}

// NNMap 3-halo, x = (log10 l1, log10 l2, phi, log10 m1..m3, z)
// (src/NNMap_Model.cu:1003-1037).
__device__ __forceinline__ double nnmap_3h(const DevTables &T, const TermArgs &a,
                                           const double *x) {
  double l1 = pow(10.0, x[0]), l2 = pow(10.0, x[1]), phi = x[2];
  double m1 = pow(10.0, x[3]), m2 = pow(10.0, x[4]), m3 = pow(10.0, x[5]), z = x[6];
  const int nb = T.nb;
  int zi = zix_round(T, z);
  int i1 = mix_round(T, m1), i2 = mix_round(T, m2), i3 = mix_round(T, m3);
  double w_ = T.w[zi];
  double cphi = cos(phi);
  double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cphi);
  double k1 = l1 / w_, k2 = l2 / w_, k3 = l3 / w_;
  double u3 = unfw(nfw_setup(T, m3, z, 1, r200(T, m3, z)), k3);
  Occ o1 = occupation(a.h1, m1);
  double G1 = o1.nc + o1.ns * unfw(nfw_setup(T, m1, z, a.h1.f, r200(T, m1, z)), k1);
  Occ o2 = occupation(a.h2, m2);
  double G2 = o2.nc + o2.ns * unfw(nfw_setup(T, m2, z, a.h2.f, r200(T, m2, z)), k2);
  double B = 1. / a.nb1[zi] / a.nb2[zi] / T.rho[zi] * T.hmf[zi * nb + i1] *
             T.hmf[zi * nb + i2] * T.hmf[zi * nb + i3] * m3 * u3 * G1 * G2 *
             bilin(T, k1, k2, cphi, z) * T.bh[zi * nb + i1] * T.bh[zi * nb + i2] *
             T.bh[zi * nb + i3];
  double b = B / T.dwdz[zi] * T.g[zi] * T.p1[zi] * T.p2[zi] / w_ / w_ / w_ * (1. + z);
  return l1 * l2 * m1 * m2 * m3 * l1 * l2 * aperture(a.t1 * l1) * aperture(a.t2 * l2) *
         aperture(a.t3 * l3) * b;
}

// NMapMap 1-halo (src/NMapMap_Model.cu:263-301).
__device__ __forceinline__ double nmapmap_1h(const DevTables &T, const TermArgs &a,
                                             const double *x) {
  double l1 = pow(10.0, x[0]), l2 = pow(10.0, x[1]), phi = x[2], m = pow(10.0, x[3]);
  double z = x[4];
  int zi = zix_round(T, z);
  int mi = mix_round(T, m);
  double nbar = a.nb1[zi];
  double w_ = T.w[zi];
  double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi));
  double k1 = l1 / w_, k2 = l2 / w_, k3 = l3 / w_;
  Occ o = occupation(a.h1, m);
  double r = r200(T, m, z);
  Nfw n1 = nfw_setup(T, m, z, 1, r);
  double uf = unfw(nfw_setup(T, m, z, a.h1.f, r), k1);
  double rho = T.rho[zi];
  double B = 1. / nbar / rho / rho * T.hmf[zi * T.nb + mi] * m * m * (o.nc + o.ns * uf) *
             unfw(n1, k2) * unfw(n1, k3);
  double b = B / T.dwdz[zi] * T.g[zi] * T.g[zi] * T.p1[zi] / w_ / w_ * (1. + z) * (1. + z);
  return l1 * l2 * m * l1 * l2 * aperture(a.t1 * l1) * aperture(a.t2 * l2) *
         aperture(a.t3 * l3) * b / nbar;
}

// NMapMap 2-halo (src/NMapMap_Model.cu:402-468).
__device__ __forceinline__ double nmapmap_2h(const DevTables &T, const TermArgs &a,
                                             const double *x) {
  double l1 = pow(10.0, x[0]), l2 = pow(10.0, x[1]), phi = x[2];
  double m1 = pow(10.0, x[3]), m2 = pow(10.0, x[4]), z = x[5];
  const int nb = T.nb;
  int zi = zix_round(T, z);
  int i1 = mix_round(T, m1), i2 = mix_round(T, m2);
  double nbar = a.nb1[zi];
  double w_ = T.w[zi];
  double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cos(phi));
  double k1 = l1 / w_, k2 = l2 / w_, k3 = l3 / w_;
  double P1 = pk(T, zi, k1), P2 = pk(T, zi, k2), P3 = pk(T, zi, k3);
  Occ o1 = occupation(a.h1, m1), o2 = occupation(a.h1, m2);
  double f = a.h1.f;
  double ra = r200(T, m1, z), rb = r200(T, m2, z);
  Nfw n1 = nfw_setup(T, m1, z, 1, ra), n2 = nfw_setup(T, m2, z, 1, rb);
  double u_k2_m1 = unfw(n1, k2), u_k3_m1 = unfw(n1, k3);
  double u_k2_m2 = unfw(n2, k2), u_k3_m2 = unfw(n2, k3);
  double u_k1_m2_f = unfw(nfw_setup(T, m2, z, f, rb), k1);
  double u_k1_m1_f = unfw(nfw_setup(T, m1, z, f, ra), k1);
  double rho = T.rho[zi];
  double B = 1. / nbar / rho / rho * T.hmf[zi * nb + i1] * T.hmf[zi * nb + i2] *
             T.bh[zi * nb + i1] * T.bh[zi * nb + i2] *
             (m1 * m2 * u_k2_m1 * u_k3_m2 * (o2.nc + o2.ns * u_k1_m2_f) * P2 +
              m1 * m2 * u_k3_m1 * u_k2_m2 * (o2.nc + o2.ns * u_k1_m2_f) * P3 +
              m2 * m2 * u_k2_m1 * u_k3_m2 * (o1.nc + o1.ns * u_k1_m1_f) * P1);
  double b = B / T.dwdz[zi] * T.g[zi] * T.g[zi] * T.p1[zi] / w_ / w_ * (1. + z) * (1. + z);
  return l1 * l2 * m1 * m2 * l1 * l2 * aperture(a.t1 * l1) * aperture(a.t2 * l2) *
         aperture(a.t3 * l3) * b / nbar;
}

// NMapMap 3-halo: Ncen, Nsat at m3 with u_NFW(k1, m2, f) (src/NMapMap_Model.cu:593-632).
__device__ __forceinline__ double nmapmap_3h(const DevTables &T, const TermArgs &a,
                                             const double *x) {
  double l1 = pow(10.0, x[0]), l2 = pow(10.0, x[1]), phi = x[2];
  double m1 = pow(10.0, x[3]), m2 = pow(10.0, x[4]), m3 = pow(10.0, x[5]), z = x[6];
  const int nb = T.nb;
  int zi = zix_round(T, z);
  int i1 = mix_round(T, m1), i2 = mix_round(T, m2), i3 = mix_round(T, m3);
  double nbar = a.nb1[zi];
  double w_ = T.w[zi];
  double cphi = cos(phi);
  double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cphi);
  double k1 = l1 / w_, k2 = l2 / w_, k3 = l3 / w_;
  Occ o = occupation(a.h1, m3);
  double ra = r200(T, m1, z), rb = r200(T, m2, z);
  double u_k2_m1 = unfw(nfw_setup(T, m1, z, 1, ra), k2);
  double u_k3_m2 = unfw(nfw_setup(T, m2, z, 1, rb), k3);
  double u_k1_m2_f = unfw(nfw_setup(T, m2, z, a.h1.f, rb), k1);
  double rho = T.rho[zi];
  double B = 1. / nbar / rho / rho * T.hmf[zi * nb + i1] * T.hmf[zi * nb + i2] *
             T.hmf[zi * nb + i3] * T.bh[zi * nb + i1] * T.bh[zi * nb + i2] *
             T.bh[zi * nb + i3] * m1 * m2 * u_k2_m1 * u_k3_m2 * (o.nc + o.ns * u_k1_m2_f) *
             bilin(T, k1, k2, cphi, z);
  double b = B / T.dwdz[zi] * T.g[zi] * T.g[zi] * T.p1[zi] / w_ / w_ * (1. + z) * (1. + z);
  return l1 * l2 * m1 * m2 * m3 * l1 * l2 * aperture(a.t1 * l1) * aperture(a.t2 * l2) *
         aperture(a.t3 * l3) * b / nbar;
}

// n_bar integrand, x = log10 m (src/generalFunctions.cu:233-244).
__device__ __forceinline__ double nbar_kernel(const DevTables &T, const DevHod &h, double z,
                                              const double *x) {
  double m = pow(10.0, x[0]);
  Occ o = occupation(h, m);
  int mi = mix_round(T, m);
  int zi = zix_round(T, z);
  return m * T.hmf[zi * T.nb + mi] * (o.nc + o.ns);
}

}  // namespace g3l
