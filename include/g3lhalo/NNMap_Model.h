// NNMap_Model.h -- <N N M_ap> from the halo model, evaluated on an MI355X.
//
// Same constructor, members and methods as the reference's
// g3lhalo::NNMap_Model (src/NNMap_Model.h:29-242); the integrals run through
// the gfx950 engine of include/g3l.h.  Single terms (NNMap_1h/2h/3h) use the
// plain form, which refines exactly like the reference's hcubature_v;
// calculateAll and the batch methods refine all integrals together on the
// device.  Errors of the engine print a message and exit, as the reference's
// CUDA_SAFE_CALL does.
#ifndef G3LHALO_NNMAP_MODEL_H
#define G3LHALO_NNMAP_MODEL_H

#include <array>
#include <vector>

#include "../g3l.h"
#include "Cosmology.h"
#include "HOD.h"
#include "constants.h"

namespace g3lhalo {

class NNMap_Model {
 public:
  Cosmology *cosmology = nullptr;
  double zmin, zmax;  // redshift binning of the tables
  double kmin, kmax;  // wave-number binning [1/Mpc]
  double mmin, mmax;  // halo-mass binning [Msun]
  int Nbins;

  // integral borders (src/NNMap_Model.cu:46-49, src/NNMap_Model.h:43-48)
  double zmin_integral, zmax_integral;
  double mmin_integral, mmax_integral;
  double phimin_integral = 0;
  double phimax_integral = 2 * g3lhalo::pi;
  double lmin_integral = 0.1;
  double lmax_integral = 1e6;

  // caller-owned tables (copied to the device at construction)
  double *g, *p_lens1, *p_lens2, *w, *dwdz, *hmf, *P_lin, *b_h, *concentration;

  HOD *hod1, *hod2;
  double A, epsilon;

  // densities per redshift bin, owned by the model
  double *rho_bar = nullptr, *n_bar1 = nullptr, *n_bar2 = nullptr;

  // Execution options (no counterpart in the reference).
  bool batched = true;           // calculateAll through g3l_eval_batch
  int precision = G3L_FP64;      // G3L_MIXED: FP32 transcendental bulk

  NNMap_Model(Cosmology *cosmology_, const double &zmin_, const double &zmax_,
              const double &kmin_, const double &kmax_, const double &mmin_,
              const double &mmax_, const int &Nbins_, double *g_, double *p_lens1_,
              double *p_lens2_, double *w_, double *dwdz_, double *hmf_, double *P_lin_,
              double *b_h_, double *concentration_, HOD *hod1_, HOD *hod2_, double A_,
              double epsilon_);
  ~NNMap_Model();
  NNMap_Model(const NNMap_Model &) = delete;
  NNMap_Model &operator=(const NNMap_Model &) = delete;

  void updateHODs(HOD *hod1_, HOD *hod2_);
  void updateDensities();

  // theta in radians; type 1 selects hod1/n_bar1, otherwise hod2/n_bar2
  double NNMap_1h(const double &theta1, const double &theta2, const double &theta3,
                  const int &type1, const int &type2);
  double NNMap_2h(const double &theta1, const double &theta2, const double &theta3,
                  const int &type1, const int &type2);
  double NNMap_3h(const double &theta1, const double &theta2, const double &theta3,
                  const int &type1, const int &type2);
  double NNMap(const double &theta1, const double &theta2, const double &theta3,
               const int &type1, const int &type2);

  // Updates the HODs, then appends NNMap for every triplet with types
  // (1,2), then (1,1), then (2,2) (src/NNMap_Model.cu:382-415).  A and
  // epsilon are unused, as in the reference (the members are used).
  void calculateAll(const std::vector<double> &thetas1, const std::vector<double> &thetas2,
                    const std::vector<double> &thetas3, HOD *hod1, HOD *hod2, double A,
                    double epsilon, std::vector<double> &results);

  // Batch (no counterpart in the reference): the 1h, 2h, 3h terms of every
  // triplet [rad] in one device-resident batch; out[i] = {1h, 2h, 3h}.
  void terms(const std::vector<std::array<double, 3>> &thetas, int type1, int type2,
             std::vector<std::array<double, 3>> &out);

 private:
  g3l_ctx *ctx_ = nullptr;
  int set_ = -1;
  double term(int t, double th1, double th2, double th3, int type1, int type2);
};

}  // namespace g3lhalo
#endif
