// doPrecomputations.x -- drop-in for the reference driver of the same name
// (src/doPrecomputations.cpp): the same 16 arguments, the same seven output
// files ("z value" rows for g, w, dw/dz; "z m value" / "z k value" rows for
// the mass function, linear power spectrum, halo bias and concentration),
// computed by g3l_precompute on the device.
#include <iostream>
#include <vector>

#include "../../../include/g3l.h"
#include "../../../include/g3lhalo/Cosmology.h"
#include "../../../include/g3lhalo/Function.h"
#include "../../../include/g3lhalo/helpers.h"

int main(int argc, char *argv[]) {
  const int n_params = 16;
  std::string usage =
      "./doPrecomputations.x \n Cosmology File \n  File with source redshift distribution \n "
      "minimal redshift \n maximal redshift \n minimal halo mass [Msun] \n maximal halo mass "
      "[Msun] \n minimal k [1/Mpc] \n maximal k [1/Mpc] Number of bins \n Filename lensing "
      "efficiency \n Filename angular diameter distance \n Filename derivative angular diameter "
      "distance \n Filename HMF \n Filename linear power spectrum \n Filename halo bias \n "
      "Filename concentration \n";
  std::string example =
      "./doPrecomputations.x cosmo.param ../results/ n_z_sources.dat 0.1 2 1e10 1e17 0.01 1000 "
      "256";
  g3lhalo::checkCmdLine(argc, n_params, usage, example);
  g3lhalo::Cosmology cosmo(argv[1]);
  g3lhalo::Function nz(argv[2], 0);
  g3l_precompute_in in{};
  in.cosmo = {cosmo.Omega_m, cosmo.Omega_Lambda, cosmo.Omega_b, cosmo.sigma8, cosmo.H0, cosmo.ns};
  in.zmin = std::stod(argv[3]);
  in.zmax = std::stod(argv[4]);
  in.mmin = std::stod(argv[5]);
  in.mmax = std::stod(argv[6]);
  in.kmin = std::stod(argv[7]);
  in.kmax = std::stod(argv[8]);
  in.nbins = std::stoi(argv[9]);
  in.nz_z = nz.x_values_.data();
  in.nz_n = nz.y_values_.data();
  in.nz_count = (int32_t)nz.x_values_.size();
  const size_t n = in.nbins > 0 ? (size_t)in.nbins : 0;
  std::vector<double> z(n), m(n), k(n), g(n), w(n), dwdz(n), hmf(n * n), plin(n * n), bh(n * n),
      conc(n * n);
  g3l_precompute_out out{z.data(), m.data(), k.data(), g.data(), w.data(), dwdz.data(),
                         hmf.data(), plin.data(), bh.data(), conc.data()};
  const int rc = g3l_precompute(&in, &out);
  if (rc != G3L_OK) {
    std::cerr << "doPrecomputations: g3l_precompute failed with status " << rc << std::endl;
    return 1;
  }
  g3lhalo::writeToFile(z, g, argv[10]);
  g3lhalo::writeToFile(z, w, argv[11]);
  g3lhalo::writeToFile(z, dwdz, argv[12]);
  g3lhalo::writeToFile(z, k, plin, argv[14]);
  g3lhalo::writeToFile(z, m, conc, argv[16]);
  g3lhalo::writeToFile(z, m, hmf, argv[13]);
  g3lhalo::writeToFile(z, m, bh, argv[15]);
  return 0;
}
