// Host build of the FP64 u_NFW pieces: reads lines "s x" or "u a1 c" (doubles
// as 16-digit hex bit patterns) on stdin and writes, for "s", sincos_mm(x) and
// sincos_cw(x), and for "u", P64::u_a1 (one common denominator, minimax
// sin/cos) and P64R::u_a1 (per-piece quotients, Taylor sin/cos) at
// c = the given concentration, F and ln(1+c) from it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../g3lhalo_amd/csrc/physics.h"

static double rd() {
  unsigned long long b;
  if (std::scanf("%llx", &b) != 1) return NAN;
  double x;
  std::memcpy(&x, &b, 8);
  return x;
}
static void wr(double y) {
  unsigned long long b;
  std::memcpy(&b, &y, 8);
  std::printf("%016llx ", b);
}

int main() {
  char op[4];
  while (std::scanf("%3s", op) == 1) {
    if (op[0] == 's') {
      const double x = rd();
      double s, c, s0, c0;
      g3l::sincos_mm(x, s, c);
      g3l::sincos_cw(x, s0, c0);
      wr(s), wr(c), wr(s0), wr(c0);
    } else {
      const double a1 = rd(), c = rd();
      g3l::Nfw n{0.0, c, std::log(1 + c), 0.0};
      n.F = n.lc - c / (1 + c);
      wr(g3l::P64::u_a1(a1, n));
      wr(g3l::P64R::u_a1(a1, n));
    }
    std::printf("\n");
  }
  return 0;
}
