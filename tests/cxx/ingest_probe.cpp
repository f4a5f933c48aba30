// Prints what the C++ API's ingest classes read from reference-format files
// (no device use): used by tests/test_cxx_ingest.py and, in its "dump" mode
// (every table exactly as the drop-in drivers resample it), by
// tests/test_gpu_pipeline.py.
#include <cstdio>
#include <string>
#include <vector>

#include "g3lhalo/Cosmology.h"
#include "g3lhalo/Function.h"
#include "g3lhalo/Function2D.h"
#include "g3lhalo/Params.h"
#include "../../g3lhalo_amd/cxx/drivers/common.h"

static void print_table(const char *name, const std::vector<double> &v) {
  std::printf("%s", name);
  for (double x : v) std::printf(" %.17g", x);
  std::printf("\n");
}

// dump Nbins zmin zmax kmin kmax mmin mmax g p1 p2 w dwdz hmf plin bh conc
static int dump(char **a) {
  drivers::Tables1D2D t;
  drivers::read_tables(std::stoi(a[0]), std::stod(a[1]), std::stod(a[2]), std::stod(a[3]),
                       std::stod(a[4]), std::stod(a[5]), std::stod(a[6]), a[7], a[8], a[9],
                       a[10], a[11], a[12], a[13], a[14], a[15], t);
  print_table("g", t.g);
  print_table("p_lens1", t.p1);
  print_table("p_lens2", t.p2);
  print_table("w", t.w);
  print_table("dwdz", t.dwdz);
  print_table("hmf", t.hmf);
  print_table("P_lin", t.plin);
  print_table("b_h", t.bh);
  print_table("concentration", t.conc);
  return 0;
}

int main(int argc, char **argv) {
  if (argc == 18 && std::string(argv[1]) == "dump") return dump(argv + 2);
  std::string dir = argv[1];
  int nb = std::stoi(argv[2]);
  double zmin = std::stod(argv[3]), zmax = std::stod(argv[4]);
  double mmin = std::stod(argv[5]), mmax = std::stod(argv[6]);
  g3lhalo::Function w(dir + "/w.dat", 0.0);
  std::vector<double> wv;
  w.read(nb, zmin, zmax, wv);
  g3lhalo::Function2D hmf(dir + "/hmf.dat", 0.0);
  std::vector<double> hv;
  hmf.read(nb, zmin, zmax, mmin, mmax, hv);
  g3lhalo::Params p(dir + "/hod1.param");
  g3lhalo::Cosmology c(dir + "/cosmo.param");
  std::printf("w");
  for (double v : wv) std::printf(" %.17g", v);
  std::printf("\nhmf");
  for (double v : hv) std::printf(" %.17g", v);
  std::printf("\nparams %.17g %.17g %.17g %.17g %.17g %.17g\n", p.f, p.alpha, p.mmin, p.sigma,
              p.mprime, p.beta);
  std::printf("cosmo %.17g %.17g %.17g %.17g\n", c.Omega_m, c.Omega_Lambda, c.H0, c.rho_crit);
  // last abscissa returns the default (src/Function.cpp:62-72)
  std::printf("edge %.17g\n", w.at(zmax));
  return 0;
}
