// select_math.h -- fixed-point error arithmetic of the batch selection
// (k_select), shared by the device kernel and host-side checks.
//
// The Gladwell batch of src/hcubature.cpp:971-985 pops regions in decreasing
// error until the remaining error meets the tolerance.  k_select finds that
// count with integer sums of the errors in fixed point, which are exact and
// independent of summation order: every error key e is mapped to
// e * 2^shift truncated to an integer.  The shift puts the largest key below
// 2^(47 - extra), where 2^(16 + extra) >= n (the number of regions), so a sum
// over all n keys stays below 2^63 and cannot wrap.  At n <= 2^16 (every
// integral whose budget is below about 1.2e7 evaluations, and every
// unlimited one, whose slot capacity is 2^16) extra is 0.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define G3L_SEL_HD __host__ __device__ inline
#else
#define G3L_SEL_HD inline
#endif

namespace g3l {

// errmax of src/hcubature.cpp:119-126 as an order-preserving integer key:
// max(0, err), NaN -> 0 (the bit pattern of a non-negative double orders as
// the double does).
G3L_SEL_HD uint64_t err_key(double e) {
  const double k = e > 0 ? e : 0.0;
  return __builtin_bit_cast(uint64_t, k);
}

// Binary exponent of the largest key, and the fixed-point shift for n keys.
G3L_SEL_HD int sel_shift(uint64_t maxkey, uint32_t n) {
  const int ex = (int)((maxkey >> 52) & 0x7ff) - 1023;
  int extra = 0;
  while (extra < 32 && (uint64_t)n > (1ull << (16 + extra))) ++extra;
  return 46 - ex - extra;
}

// A key in fixed-point units.
G3L_SEL_HD unsigned long long fixed_err(uint64_t k, int shift) {
  return (unsigned long long)ldexp(__builtin_bit_cast(double, k), shift);
}

// Error (in fixed-point units) the batch must remove: the total minus the
// tolerance |V| reltol, at least one unit (at least one region is popped).
G3L_SEL_HD uint64_t batch_need(uint64_t total, double thr) {
  if (thr >= (double)total) return 1;
  const uint64_t need = total - (uint64_t)floor(thr);
  return need == 0 ? 1 : need;
}

}  // namespace g3l
