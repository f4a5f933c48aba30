// Function2D.h -- 2-D table ingest (the reference's g3lhalo::Function2D,
// src/Function2D.h): "x1 x2 y" triples of a square grid, nearest-index
// look-up linear in x1 and logarithmic in x2 (src/Function2D.cpp:46-64).
// read() re-targets the look-up to the model grid before sampling it, so a
// file written on the model grid is returned entry for entry
// (src/Function2D.cpp:66-95).
#ifndef G3LHALO_FUNCTION2D_H
#define G3LHALO_FUNCTION2D_H

#include <string>
#include <vector>

namespace g3lhalo {
class Function2D {
 public:
  Function2D() {}
  Function2D(std::string filename, const double &default_value);
  double at(const double &x1, const double &x2);
  int read(const int &Nbins, const double &min1, const double &max1, const double &min2,
           const double &max2, std::vector<double> &values);

 private:
  std::vector<double> x1_values_, x2_values_, y_values_;
  double default_value_ = 0;
  double x1_min = 0, x1_max = 0, x2_min = 0, x2_max = 0;
  int N1 = 0, N2 = 0;
};
}  // namespace g3lhalo
#endif
