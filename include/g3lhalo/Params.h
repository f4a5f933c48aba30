// Params.h -- HOD parameters (the reference's g3lhalo::Params,
// src/Params.h:15-46): f, alpha, M_min (threshold mass M_th), sigma, M', beta.
#ifndef G3LHALO_PARAMS_H
#define G3LHALO_PARAMS_H

#include <ostream>
#include <string>

namespace g3lhalo {
class Params {
 public:
  double f = 0;       // satellite concentration factor
  double alpha = 0;   // central amplitude
  double mmin = 0;    // threshold mass [Msun]
  double sigma = 0;   // width of the central step
  double mprime = 0;  // satellite mass scale [Msun]
  double beta = 0;    // satellite slope

  Params() {}
  Params(double f_, double alpha_, double mmin_, double sigma_, double mprime_, double beta_)
      : f{f_}, alpha{alpha_}, mmin{mmin_}, sigma{sigma_}, mprime{mprime_}, beta{beta_} {}
  // "name value" lines, '#' comments; names M_min, sigma, M', beta, alpha, f.
  // An unknown name or unreadable file prints a message and exits(1), as the
  // reference does (src/Params.cpp:7-74).
  explicit Params(const std::string &filename);
};
}  // namespace g3lhalo

std::ostream &operator<<(std::ostream &out, g3lhalo::Params const &params);
#endif
