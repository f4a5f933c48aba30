// Host check of the batch selection's fixed-point error arithmetic
// (g3lhalo_amd/csrc/select_math.h): for n equal error keys the fixed-point
// total must equal n times one key exactly (no wrap past 2^64), and the
// number of pops that the kernel derives from it must equal the exact count
// ceil-based count of src/hcubature.cpp:971-985's loop, n - floor(thr / e).
// Prints "ok" or the first failure.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <initializer_list>

#include "../../g3lhalo_amd/csrc/select_math.h"

using namespace g3l;

static int check(uint32_t n, double e, double V, double tol) {
  const uint64_t k = err_key(e);
  const int shift = sel_shift(k, n);
  const unsigned long long fx = fixed_err(k, shift);
  const unsigned __int128 total = (unsigned __int128)fx * n;
  if (total >> 63) {
    std::printf("n=%u e=%g: fixed total %llu * %u reaches 2^63\n", n, e, fx, n);
    return 1;
  }
  const uint64_t tot = (uint64_t)total;
  const double thr = std::ldexp(std::fabs(V) * tol, shift);
  const uint64_t need = batch_need(tot, thr);
  // all keys tie: the kernel takes r = ceil(need / fx) of them (at least one)
  uint64_t r = (need + fx - 1) / fx;
  if (r < 1) r = 1;
  if (r > n) r = n;
  // the reference loop: pop until (n - K) e <= |V| tol, at least once
  uint64_t K = 0;
  double rem = (double)n * e;
  const double lim = std::fabs(V) * tol;
  do {
    ++K;
    rem = (double)(n - K) * e;
  } while (K < n && rem > lim);
  if (r != K) {
    std::printf("n=%u e=%g V=%g: kernel count %llu, reference %llu\n", n, e, V,
                (unsigned long long)r, (unsigned long long)K);
    return 1;
  }
  return 0;
}

int main() {
  int bad = 0;
  // n around and beyond 2^16, 2^17 and up to the 2^26 slot capacity
  const uint32_t ns[] = {1, 2, 1000, 65535, 65536, 65537, 131072, 131077, 262145, 1u << 20,
                         (1u << 26) - 1, 1u << 26};
  const double es[] = {1.0, 0.75, 3.0e-7, 1.9e5, 1e-300};
  for (uint32_t n : ns)
    for (double e : es)
      for (double frac : {0.0, 0.3, 0.5, 0.999}) {
        // |V| tol is frac of the total error, so about (1 - frac) n regions pop
        const double V = frac * n * e / 1e-4;
        bad += check(n, e, V, 1e-4);
      }
  // the shift for n <= 2^16 is the unscaled one (production budgets keep it)
  if (sel_shift(err_key(1.0), 65536) != 46 || sel_shift(err_key(1.0), 65537) != 45) {
    std::printf("shift for n <= 2^16 changed\n");
    ++bad;
  }
  if (!bad) std::printf("ok\n");
  return bad ? 1 : 0;
}
