// Host build of g3l::log_pos (the device's ln x of Ci(x) in u_NFW): reads
// doubles as 16-digit hex bit patterns on stdin, writes log_pos(x) the same way.
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../g3lhalo_amd/csrc/physics.h"

int main() {
  unsigned long long b;
  while (std::scanf("%llx", &b) == 1) {
    double x;
    std::memcpy(&x, &b, 8);
    const double y = g3l::log_pos(x);
    std::memcpy(&b, &y, 8);
    std::printf("%016llx\n", b);
  }
  return 0;
}
