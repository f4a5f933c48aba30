// HOD.h -- halo occupation distribution (the reference's g3lhalo::HOD,
// src/HOD.h:22-59).  Host evaluations of the occupation functions; the
// integrals evaluate them on the device.
#ifndef G3LHALO_HOD_H
#define G3LHALO_HOD_H

#include "Params.h"

namespace g3lhalo {
class HOD {
 public:
  Params *params = nullptr;
  HOD() {}
  explicit HOD(Params *params_) : params{params_} {}
  // 0.5 (1 + erf(ln(m/M_th)/sigma/1.414213562)) (m/M')^beta (src/HOD.cu:35-56)
  double Nsat(double m, double *unused = nullptr) const;
  // 0.5 alpha (1 + erf(ln(m/M_th)/sigma/1.414213562)) (src/HOD.cu:58-76)
  double Ncen(double m, double *unused = nullptr) const;
};
// Same object: (Ns + scale1^2 - 1) Ns; otherwise max(0, Ns1 Ns2 + A m^eps
// sqrt(Ns1 Ns2)) (src/HOD.cu:78-94).
double NsatNsat(double m, HOD *hod1, HOD *hod2, double *d_params1 = nullptr,
                double *d_params2 = nullptr, double A = 0, double epsilon = 0,
                double scale1 = 1, double scale2 = 1);
}  // namespace g3lhalo
#endif
