// calculateNMapMap.x -- drop-in for the reference driver of the same name
// (src/calculateNMapMap.cu): same 18 arguments, same stdout rows
// "theta1 theta2 theta3 1h 2h 3h".
#include "../../../include/g3lhalo/Cosmology.h"
#include "../../../include/g3lhalo/HOD.h"
#include "../../../include/g3lhalo/NMapMap_Model.h"
#include "../../../include/g3lhalo/Params.h"
#include "../../../include/g3lhalo/helpers.h"
#include "common.h"

int main(int argc, char *argv[]) {
  const int n_params = 18;
  std::string usage =
      "./calculateNMapMap.x \n Cosmology File \n minimal redshift \n maximal redshift \n minimal "
      "k [1/Mpc] \n maximal k [1/Mpc] \n minimal halo mass [Msun] \n maximal halo mass [Msun] \n "
      "Number of bins \n File for lensing efficiency \n File for lens redshift distribution \n "
      "File for comoving distance \n File for derivative of comoving distance \n File for Halo "
      "Mass function \n File for Linear power spectrum \n File for Halo bias \n File for "
      "concentration \n File for Halomodel Params \n File with thetas";
  std::string example =
      "./calculateNMapMap.x cosmo.param 0.1 2 0.01 1000 1e10 1e17 256 g.dat plens.dat w.dat "
      "dwdz.dat hmf.dat Plin.dat b_h.dat conc.dat hod.param thetas.dat";
  g3lhalo::checkCmdLine(argc, n_params, usage, example);
  std::string fn_cosmo = argv[1];
  double zmin = std::stod(argv[2]), zmax = std::stod(argv[3]);
  double kmin = std::stod(argv[4]), kmax = std::stod(argv[5]);
  double mmin = std::stod(argv[6]), mmax = std::stod(argv[7]);
  int Nbins = std::stoi(argv[8]);
  drivers::Tables1D2D t;
  drivers::read_tables(Nbins, zmin, zmax, kmin, kmax, mmin, mmax, argv[9], argv[10], argv[10],
                       argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], t);
  g3lhalo::Params params(argv[17]);
  g3lhalo::HOD hod(&params);
  g3lhalo::Cosmology cosmo(fn_cosmo);
  g3lhalo::NMapMap_Model nmapmap(&cosmo, zmin, zmax, kmin, kmax, mmin, mmax, Nbins, t.g.data(),
                                 t.p1.data(), t.w.data(), t.dwdz.data(), t.hmf.data(),
                                 t.plin.data(), t.bh.data(), t.conc.data(), &hod);
  std::vector<std::array<double, 3>> arcmin, terms;
  drivers::read_thetas(argv[18], arcmin);
  nmapmap.terms(drivers::to_rad(arcmin), terms);
  drivers::print_rows(arcmin, terms);
  return 0;
}
