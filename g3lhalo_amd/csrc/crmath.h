// crmath.h -- correctly rounded 10^x and cos(x) in double-double arithmetic.
//
// The integrands' coordinates l = 10^x and cos(phi) feed the tree-level
// bispectrum, whose squeezed configurations (k3 -> 0, where the cosine-rule
// angles lose most of their digits) amplify a last-place difference in them
// by many orders of magnitude.  The reference takes them from the platform's
// pow and cos, which in the host library it was built against (glibc) return
// the correctly rounded result for all but about 0.1 % of arguments.  These
// routines return the correctly rounded result (the double nearest the exact
// value; error bound of the double-double evaluation below 2^-100 relative),
// so the device's coordinates equal the reference build's wherever that
// library rounds correctly.
//
// Double-double arithmetic after Dekker (1971) and Knuth's two-sum: a value
// is hi + lo with |lo| <= ulp(hi)/2; products from one FMA.  Series: Taylor
// polynomials evaluated by Horner's rule, the high-order terms (below 2^-55
// relative) in plain double, the rest in double-double.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define G3L_CR_HD __host__ __device__ __forceinline__
#else
#define G3L_CR_HD inline
#endif

namespace g3l {
namespace cr {

struct DD {
  double hi, lo;
};

G3L_CR_HD DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return DD{s, (a - (s - bb)) + (b - bb)};
}
G3L_CR_HD DD fast_two_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return DD{s, b - (s - a)};
}
G3L_CR_HD DD two_prod(double a, double b) {
  const double p = a * b;
  return DD{p, fma(a, b, -p)};
}
G3L_CR_HD DD mul(DD a, DD b) {
  DD p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}
G3L_CR_HD DD mul_d(DD a, double b) {
  DD p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}
G3L_CR_HD DD add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  s.lo += a.lo + b.lo;
  return fast_two_sum(s.hi, s.lo);
}

// 1/j! as double-double (hi + lo), j = 0..27
constexpr double kIFH[28] = {1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333, 0.001388888888888889, 0.0001984126984126984, 2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07, 2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10, 1.1470745597729725e-11, 7.647163731819816e-13, 4.779477332387385e-14, 2.8114572543455206e-15, 1.5619206968586225e-16, 8.22063524662433e-18, 4.110317623312165e-19, 1.9572941063391263e-20, 8.896791392450574e-22, 3.868170170630684e-23, 1.6117375710961184e-24, 6.446950284384474e-26, 2.4795962632247976e-27, 9.183689863795546e-29};
constexpr double kIFL[28] = {0.0, 0.0, 0.0, 9.25185853854297e-18, 2.3129646346357427e-18, 1.1564823173178714e-19, -5.300543954373577e-20, 1.7209558293420705e-22, 2.1511947866775882e-23, -1.858393274046472e-22, 2.3767714622250297e-23, -1.448814070935912e-24, -1.20734505911326e-25, 1.2585294588752098e-26, 2.0655512752830745e-28, 7.03872877733453e-30, 4.399205485834081e-31, 1.6508842730861433e-31, 1.1910679660273754e-32, 2.2141894119604265e-34, 1.4412973378659527e-36, -1.3643503830087908e-36, -7.911402614872376e-38, -8.843177655482344e-40, -3.6846573564509766e-41, -1.9330404233703465e-42, -1.2953730964765229e-43, 1.4303150396787322e-45};

// exp(a) for |a| <= 0.05 as double-double: Taylor through a^14 (the first
// omitted term is below 2^-110 relative); the terms from a^9 on sum to below
// 2^-58 relative, so their double value suffices.
G3L_CR_HD DD exp_small(DD a) {
  double t = kIFH[14];
#pragma unroll
  for (int k = 13; k >= 9; --k) t = fma(t, a.hi, kIFH[k]);
  DD acc = add(two_prod(t, a.hi), DD{kIFH[8], kIFL[8]});
#pragma unroll
  for (int k = 7; k >= 0; --k) acc = add(mul(acc, a), DD{kIFH[k], kIFL[k]});
  return acc;
}

// 10^x, correctly rounded, for |x| < 300.
G3L_CR_HD double exp10_cr(double x) {
  // x log2(10) = n + r, |r| <= 1/2 + tiny, in double-double
  const DD p = two_prod(x, 3.321928094887362);
  const double tlo = fma(x, 1.661617516973592e-16, p.lo);
  const double n = rint(p.hi);
  const DD r = two_sum(p.hi - n, tlo);  // p.hi - n is exact (Sterbenz)
  // 2^r = exp(r ln2) = exp(r ln2 / 8)^8
  DD s = mul(r, DD{0.6931471805599453, 2.3190468138462996e-17});
  s.hi *= 0.125;
  s.lo *= 0.125;
  DD e = exp_small(s);
  e = mul(e, e);
  e = mul(e, e);
  e = mul(e, e);
  return ldexp(e.hi + e.lo, (int)n);
}

// cos(x), correctly rounded, for 0 <= x <= 8.
G3L_CR_HD double cos_cr(double x) {
  // x = n pi/2 + r, |r| <= pi/4 + tiny: pi/2 = p1 + p2 + p3 with p1, p2 of
  // 33 significant bits (n p1, n p2 exact for n < 2^20)
  const double p1 = 1.5707963267341256, p2 = 6.077100506303966e-11, p3 = 2.0222662487959506e-21;
  const double n = rint(x * 0.6366197723675814);
  const double r1 = x - n * p1;           // exact: n p1 is exact and within a factor 2 of x
  DD r = two_sum(r1, -(n * p2));          // n p2 exact
  r.lo -= n * p3;
  r = fast_two_sum(r.hi, r.lo);
  const int q = (int)n & 3;
  // cos r (q even) or sin r (q odd) = sum_k (-1)^k r^(2k+o)/(2k+o)!, o = q & 1,
  // through k = 13 (the first omitted term is below 2^-110 at |r| = pi/4);
  // the terms with 2k + o >= 18 (below 2^-56 relative) in plain double
  const int o = q & 1;
  const DD u = mul(r, r);
  double t = 0.0;
#pragma unroll
  for (int k = 13; k >= 9; --k) {
    const double c = o ? kIFH[2 * k + 1] : kIFH[2 * k];
    t = fma(t, u.hi, (k & 1) ? -c : c);
  }
  DD acc{t, 0.0};  // the Horner value at k = 9
#pragma unroll
  for (int k = 8; k >= 0; --k) {
    const double sg = (k & 1) ? -1.0 : 1.0;
    const double ch = o ? kIFH[2 * k + 1] : kIFH[2 * k];
    const double cl = o ? kIFL[2 * k + 1] : kIFL[2 * k];
    acc = add(mul(acc, u), DD{sg * ch, sg * cl});
  }
  if (o) acc = mul(acc, r);
  const double v = acc.hi + acc.lo;
  return (q == 1 || q == 2) ? -v : v;
}

}  // namespace cr
}  // namespace g3l
