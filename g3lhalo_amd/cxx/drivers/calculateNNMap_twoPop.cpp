// calculateNNMap_twoPop.x -- drop-in for the reference driver of the same
// name (src/calculateNNMap_twoPop.cu): the same 22 arguments and three
// blocks of rows, N1N2 (types 1,2), N1N1 (1,1), N2N2 (2,2).
#include "../../../include/g3lhalo/Cosmology.h"
#include "../../../include/g3lhalo/HOD.h"
#include "../../../include/g3lhalo/NNMap_Model.h"
#include "../../../include/g3lhalo/Params.h"
#include "../../../include/g3lhalo/helpers.h"
#include "common.h"

int main(int argc, char *argv[]) {
  const int n_params = 22;
  std::string usage =
      "./calculateNNMap_twoPop.x \n Cosmology File \n minimal redshift \n maximal redshift \n "
      "minimal k [1/Mpc] \n maximal k [1/Mpc] \n minimal halo mass [Msun] \n maximal halo mass "
      "[Msun] \n Number of bins \n File for lensing efficiency \n File for lens redshift "
      "distribution 1 \n File for lens redshift distribution 2 \n File for comoving distance \n "
      "File for derivative of comoving distance \n File for Halo Mass function \n File for "
      "Linear power spectrum \n File for Halo bias \n File for concentration \n File for "
      "Halomodel Params 1 \n File for Halomodel Params 2 \n Parameter A \n Parameter epsilon \n "
      "File with thetas";
  std::string example =
      "./calculateNNMap_twoPop.x cosmo.param 0.1 2 0.01 1000 1e10 1e17 256 g.dat plens1.dat "
      "plens2.dat w.dat dwdz.dat hmf.dat Plin.dat b_h.dat conc.dat hod1.param hod2.param 0 0 "
      "thetas.dat";
  g3lhalo::checkCmdLine(argc, n_params, usage, example);
  std::string fn_cosmo = argv[1];
  double zmin = std::stod(argv[2]), zmax = std::stod(argv[3]);
  double kmin = std::stod(argv[4]), kmax = std::stod(argv[5]);
  double mmin = std::stod(argv[6]), mmax = std::stod(argv[7]);
  int Nbins = std::stoi(argv[8]);
  drivers::Tables1D2D t;
  drivers::read_tables(Nbins, zmin, zmax, kmin, kmax, mmin, mmax, argv[9], argv[10], argv[11],
                       argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], t);
  g3lhalo::Params params1(argv[18]), params2(argv[19]);
  g3lhalo::HOD hod1(&params1), hod2(&params2);
  double A = std::stod(argv[20]), epsilon = std::stod(argv[21]);
  g3lhalo::Cosmology cosmo(fn_cosmo);
  g3lhalo::NNMap_Model nnmap(&cosmo, zmin, zmax, kmin, kmax, mmin, mmax, Nbins, t.g.data(),
                             t.p1.data(), t.p2.data(), t.w.data(), t.dwdz.data(), t.hmf.data(),
                             t.plin.data(), t.bh.data(), t.conc.data(), &hod1, &hod2, A,
                             epsilon);
  std::vector<std::array<double, 3>> arcmin, terms;
  drivers::read_thetas(argv[22], arcmin);
  const auto rad = drivers::to_rad(arcmin);
  const int types[3][2] = {{1, 2}, {1, 1}, {2, 2}};
  for (auto &ty : types) {
    nnmap.terms(rad, ty[0], ty[1], terms);
    drivers::print_rows(arcmin, terms);
  }
  return 0;
}
