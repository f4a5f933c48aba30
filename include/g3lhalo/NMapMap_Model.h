// NMapMap_Model.h -- <N M_ap M_ap> from the halo model, evaluated on an
// MI355X.  Same constructor, members and methods as the reference's
// g3lhalo::NMapMap_Model (src/NMapMap_Model.h:28-170); see NNMap_Model.h.
#ifndef G3LHALO_NMAPMAP_MODEL_H
#define G3LHALO_NMAPMAP_MODEL_H

#include <array>
#include <vector>

#include "../g3l.h"
#include "Cosmology.h"
#include "HOD.h"
#include "constants.h"

namespace g3lhalo {

class NMapMap_Model {
 public:
  Cosmology *cosmology = nullptr;
  double zmin, zmax, kmin, kmax, mmin, mmax;
  int Nbins;
  double zmin_integral, zmax_integral, mmin_integral, mmax_integral;
  double phimin_integral = 0;
  double phimax_integral = 2 * g3lhalo::pi;
  double lmin_integral = 0.1;
  double lmax_integral = 1e6;
  double *g, *p_lens, *w, *dwdz, *hmf, *P_lin, *b_h, *concentration;
  HOD *hod;
  double *rho_bar = nullptr, *n_bar = nullptr;

  bool batched = true;
  int precision = G3L_FP64;

  NMapMap_Model(Cosmology *cosmology_, const double &zmin_, const double &zmax_,
                const double &kmin_, const double &kmax_, const double &mmin_,
                const double &mmax_, const int &Nbins_, double *g_, double *p_lens_, double *w_,
                double *dwdz_, double *hmf_, double *P_lin_, double *b_h_,
                double *concentration_, HOD *hod_);
  ~NMapMap_Model();
  NMapMap_Model(const NMapMap_Model &) = delete;
  NMapMap_Model &operator=(const NMapMap_Model &) = delete;

  void updateHOD(HOD *hod_);
  void updateDensities();

  double NMapMap_1h(const double &theta1, const double &theta2, const double &theta3);
  double NMapMap_2h(const double &theta1, const double &theta2, const double &theta3);
  double NMapMap_3h(const double &theta1, const double &theta2, const double &theta3);
  double NMapMap(const double &theta1, const double &theta2, const double &theta3);

  // Updates the HOD, then appends NMapMap for every triplet
  // (src/NMapMap_Model.cu:247-261).
  void calculateAll(const std::vector<double> &thetas1, const std::vector<double> &thetas2,
                    const std::vector<double> &thetas3, HOD *hod_,
                    std::vector<double> &results);

  void terms(const std::vector<std::array<double, 3>> &thetas,
             std::vector<std::array<double, 3>> &out);

 private:
  g3l_ctx *ctx_ = nullptr;
  int set_ = -1;
  double term(int t, double th1, double th2, double th3);
};

}  // namespace g3lhalo
#endif
