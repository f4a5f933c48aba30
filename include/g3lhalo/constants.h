// constants.h -- numerical constants of the reference, with its truncations
// (src/constants.h:15-42): pi = 3.14159 and 2.909e-4 rad per arcmin are part
// of the results' parity.
#ifndef G3LHALO_CONSTANTS_H
#define G3LHALO_CONSTANTS_H

#include <cmath>

namespace g3lhalo {
const double pi = 3.14159;
const double sq2 = std::sqrt(2.0);
const double ln10 = std::log(10.0);
const double solarmass_in_kg = 1.98845e30;
const double Mpc_in_m = 3.086e22;
const double rad_in_arcmin = pi / 180.0 / 60.0;
const double arcmin_in_rad = 2.909e-4;
const double c = 2.99792e5;  // speed of light [km/s]
const double G = 6.674e-11;  // [m^3/s^2/kg]
}  // namespace g3lhalo
#endif
