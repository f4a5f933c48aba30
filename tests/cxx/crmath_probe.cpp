// Host driver of the correctly rounded 10^x and cos(x) of
// g3lhalo_amd/csrc/crmath.h (the same source the device compiles): reads
// lines "e x" (10^x) or "c x" (cos x) with x as a hexadecimal float and
// prints each result as a hexadecimal float.
#include <cstdio>
#include <cstdlib>

#include "../../g3lhalo_amd/csrc/crmath.h"

int main() {
  char op; double x;
  while (scanf(" %c %la", &op, &x) == 2) {
    double r = op == 'e' ? g3l::cr::exp10_cr(x) : g3l::cr::cos_cr(x);
    printf("%a\n", r);
  }
}
