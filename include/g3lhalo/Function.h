// Function.h -- 1-D table ingest (the reference's g3lhalo::Function,
// src/Function.h): "x y" pairs, linear interpolation inside [x0, xN), the
// default value outside and at x = xN exactly (src/Function.cpp:52-73).
#ifndef G3LHALO_FUNCTION_H
#define G3LHALO_FUNCTION_H

#include <string>
#include <vector>

namespace g3lhalo {
class Function {
 public:
  Function() {}
  Function(std::string filename, const double &default_value);
  double at(const double &x);
  double x_lower(const double &y);
  // values at x_i = min + i (max - min)/Nbins, i < Nbins (src/Function.cpp:89-105)
  int read(const int &Nbins, const double &min, const double &max, std::vector<double> &values);
  // values at x_i = exp(ln min + i ln(max/min)/Nbins) (src/Function.cpp:108-124)
  int readLog(const int &Nbins, const double &min, const double &max,
              std::vector<double> &values);

  // the samples as read, public as in the reference (src/Function.h), which
  // doPrecomputations passes on directly (src/doPrecomputations.cpp:55-62)
  std::vector<double> x_values_, y_values_;

 private:
  double default_value_ = 0;
};
}  // namespace g3lhalo
#endif
