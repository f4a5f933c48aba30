// rules.h -- cubature rules on the device: point q of a region and the
// per-region reduction to (value, error, split dimension).
//
// Genz-Malik 7/5 (dimension >= 2) and Gauss-Kronrod 7/15 (dimension 1), with
// the point order, weights and summation order of src/hcubature.cpp:343-614
// and :620-765, so that equal integrand values give equal bits.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

#pragma clang fp contract(off)

namespace g3l {

constexpr double kLam2 = 0.3585685828003180919906451539079374954541;  // sqrt(9/70)
constexpr double kLam4 = 0.9486832980505137995996680633298155601160;  // sqrt(9/10)
constexpr double kLam5 = 0.6882472016116852977216287342936235251269;  // sqrt(9/19)

__constant__ constexpr double kXgk[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144838258730, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
__constant__ constexpr double kWg[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
__constant__ constexpr double kWgk[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};

// Pair index -> (i, j), i < j, in the order i = 0.., j = i+1.. of
// src/hcubature.cpp:370-390.
__device__ __forceinline__ void pair_of(int D, int p, int &i, int &j) {
  int a = 0, rem = p;
  while (rem >= D - 1 - a) {
    rem -= D - 1 - a;
    ++a;
  }
  i = a;
  j = a + 1 + rem;
}

// Coordinates of point q of a region with centre c and half-width h.
template <int D>
__device__ __forceinline__ void rule_point(int q, const double *c, const double *h, double *x) {
  if constexpr (D == 1) {
    double w;
    if (q == 0) {
      x[0] = c[0];
      return;
    }
    int r = q - 1;          // 0..13
    int j = r >> 1;         // 0..6
    int j2 = j < 3 ? 2 * j + 1 : 2 * (j - 3);
    w = h[0] * kXgk[j2];
    x[0] = (r & 1) ? c[0] + w : c[0] - w;
    return;
  }
#pragma unroll
  for (int i = 0; i < D; ++i) x[i] = c[i];
  if (q == 0) return;
  int t = q - 1;
  if (t < 4 * D) {
    int k = t >> 2, s = t & 3;
    double r = h[k] * ((s < 2) ? kLam2 : kLam4);
    x[k] = (s & 1) ? c[k] + r : c[k] - r;
    return;
  }
  t -= 4 * D;
  if (t < 2 * D * (D - 1)) {
    int p = t >> 2, s = t & 3, i, j;
    pair_of(D, p, i, j);
    double ri = h[i] * kLam4, rj = h[j] * kLam4;
    x[i] = (s == 1 || s == 2) ? c[i] + ri : c[i] - ri;
    x[j] = (s >= 2) ? c[j] + rj : c[j] - rj;
    return;
  }
  t -= 2 * D * (D - 1);
  unsigned signs = (unsigned)t ^ ((unsigned)t >> 1);
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double r = h[i] * kLam5;
    x[i] = ((signs >> i) & 1u) ? c[i] - r : c[i] + r;
  }
}

// s + p[0] + p[1] + ... + p[N-1], added in that order.  The values are
// loaded in batches of 16 before they are added: left to the compiler, the
// chain kept one LDS load in flight and waited a full LDS round trip for
// every two additions.
template <int N>
__device__ __forceinline__ double ordered_sum(const double *p, double s) {
  constexpr int B = 16;
#pragma unroll
  for (int k0 = 0; k0 < N; k0 += B) {
    double buf[B];
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (k0 + j < N) buf[j] = p[k0 + j];
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (k0 + j < N) s += buf[j];
  }
  return s;
}

// The sum of the 4th (2 D (D - 1) points, fifth false) or 5th (2^D points)
// group of one region's P values, in the order of gm_reduce.  Both start with
// their first min(2 D (D - 1), 2^D) terms, so that two lanes of a wave can
// take the two sums of one region in one instruction stream; the longer
// sum's lane then adds the rest alone.
template <int D>
__device__ __forceinline__ double gm_group_sum(const double *v, bool fifth) {
  constexpr int k4 = 1 + 4 * D, n4 = 2 * D * (D - 1), n5 = 1 << D, k5 = k4 + n4;
  constexpr int nc = n4 < n5 ? n4 : n5;
  const double *p = v + (fifth ? k5 : k4);
  double s = ordered_sum<nc>(p, 0.0);
  if constexpr (n5 > n4) {
    if (fifth) s = ordered_sum<n5 - nc>(p + nc, s);
  } else if constexpr (n4 > n5) {
    if (!fifth) s = ordered_sum<n4 - nc>(p + nc, s);
  }
  return s;
}

// Genz-Malik reduction of one region's P values (src/hcubature.cpp:512-577),
// given the sums s4, s5 of its 4th and 5th point groups (gm_group_sum).
template <int D>
__device__ __forceinline__ void gm_finish(const double *v, double vol, double s4, double s5,
                                          double &val, double &err, int &split) {
  const double w2 = 980. / 6561., w4 = 200. / 19683.;
  const double e2 = 245. / 486., e4 = 25. / 729.;
  const double ratio = (kLam2 * kLam2) / (kLam4 * kLam4);
  const double w1 = (double)(12824 - 9120 * D + 400 * D * D) / (double)19683;
  const double w3 = (double)(1820 - 400 * D) / (double)19683;
  const double w5 = (double)6859 / (double)19683 / (double)(1u << D);
  const double e1 = (double)(729 - 950 * D + 50 * D * D) / (double)729;
  const double e3 = (double)(265 - 100 * D) / (double)1458;
  double f0 = v[0];
  double s2 = 0, s3 = 0;
  double best = 0;
  int arg = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double a = v[1 + 4 * k], b = v[2 + 4 * k], c = v[3 + 4 * k], e = v[4 + 4 * k];
    s2 += a + b;
    s3 += c + e;
    double diff = fabs(a + b - 2 * f0 - ratio * (c + e - 2 * f0));
    if (diff > best) {
      best = diff;
      arg = k;
    }
  }
  double r7 = vol * (w1 * f0 + w2 * s2 + w3 * s3 + w4 * s4 + w5 * s5);
  double r5 = vol * (e1 * f0 + e2 * s2 + e3 * s3 + e4 * s4);
  val = r7;
  err = fabs(r5 - r7);
  split = arg;
}
template <int D>
__device__ __forceinline__ void gm_reduce(const double *v, double vol, double &val, double &err,
                                          int &split) {
  gm_finish<D>(v, vol, gm_group_sum<D>(v, false), gm_group_sum<D>(v, true), val, err, split);
}

// Gauss-Kronrod 7/15 with the QUADPACK error scaling (src/hcubature.cpp:695-757).
__device__ __forceinline__ void gk_reduce(const double *v, double h, double &val, double &err) {
  double rg = v[0] * kWg[3];
  double rk = v[0] * kWgk[7];
  double rabs = fabs(rk);
  int n = 1;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    int j2 = 2 * j + 1;
    double s = v[n] + v[n + 1];
    rg += kWg[j] * s;
    rk += kWgk[j2] * s;
    rabs += kWgk[j2] * (fabs(v[n]) + fabs(v[n + 1]));
    n += 2;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int j2 = 2 * j;
    rk += kWgk[j2] * (v[n] + v[n + 1]);
    rabs += kWgk[j2] * (fabs(v[n]) + fabs(v[n + 1]));
    n += 2;
  }
  val = rk * h;
  double mean = rk * 0.5;
  double rasc = kWgk[7] * fabs(v[0] - mean);
  n = 1;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    int j2 = 2 * j + 1;
    rasc += kWgk[j2] * (fabs(v[n] - mean) + fabs(v[n + 1] - mean));
    n += 2;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int j2 = 2 * j;
    rasc += kWgk[j2] * (fabs(v[n] - mean) + fabs(v[n + 1] - mean));
    n += 2;
  }
  double e = fabs(rk - rg) * h;
  rabs *= h;
  rasc *= h;
  if (rasc != 0 && e != 0) {
    double scale = pow((200 * e / rasc), 1.5);
    e = (scale < 1) ? rasc * scale : rasc;
  }
  const double dbl_min = 2.2250738585072014e-308, dbl_eps = 2.220446049250313e-16;
  if (rabs > dbl_min / (50 * dbl_eps)) {
    double min_err = 50 * dbl_eps * rabs;
    if (min_err > e) e = min_err;
  }
  err = e;
}

}  // namespace g3l
