// Prints what the C++ API's ingest classes read from reference-format files
// (no device use): used by tests/test_cxx_ingest.py.
#include <cstdio>
#include <string>
#include <vector>

#include "g3lhalo/Cosmology.h"
#include "g3lhalo/Function.h"
#include "g3lhalo/Function2D.h"
#include "g3lhalo/Params.h"

int main(int argc, char **argv) {
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
