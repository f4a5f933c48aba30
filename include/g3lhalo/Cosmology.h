// Cosmology.h -- the cosmological parameters the integration path reads
// (the reference's g3lhalo::Cosmology, src/Cosmology.h:14-60).
#ifndef G3LHALO_COSMOLOGY_H
#define G3LHALO_COSMOLOGY_H

#include <string>

#include "constants.h"

namespace g3lhalo {
class Cosmology {
 public:
  double Omega_m = 0, Omega_Lambda = 0, Omega_b = 0, sigma8 = 0, H0 = 0, ns = 0;
  double c = 299792;        // [km/s]
  double G = 1.327e11;      // [km^3/Msun/s^2]
  double Mpc_in_km = 3.086e19;
  double rho_crit = 0;      // 3 H0^2 / (8 pi G) [Msun/Mpc^3]

  Cosmology() {}
  Cosmology(const double &Omega_m_, const double &Omega_Lambda_, const double &Omega_b_,
            const double &sigma8_, const double &H0_, const double &ns_)
      : Omega_m{Omega_m_}, Omega_Lambda{Omega_Lambda_}, Omega_b{Omega_b_}, sigma8{sigma8_},
        H0{H0_}, ns{ns_} {
    rho_crit = 3 * H0 * H0 * Mpc_in_km / 8 / g3lhalo::pi / G;
  }
  // "name value" lines with Omega_m, Omega_Lambda, Omega_b, sigma8, H0, ns
  // (src/Cosmology.cpp:11-72); an unknown name prints a message and stops
  // reading, leaving rho_crit unset, as the reference does.
  explicit Cosmology(const std::string &fn_parameters);

  double E(const double &z) const;         // Omega_m (1+z)^3 + Omega_Lambda
  double Omega_m_(const double &z) const;  // Omega_m (1+z)^3 / E(z)
};
}  // namespace g3lhalo
#endif
