// Shared argument handling of the drop-in drivers (reference:
// src/calculateNNMap.cu:14-129, src/calculateNMapMap.cu:14-120).
#ifndef G3LHALO_DRIVERS_COMMON_H
#define G3LHALO_DRIVERS_COMMON_H

#include <array>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../../include/g3lhalo/Function.h"
#include "../../../include/g3lhalo/Function2D.h"
#include "../../../include/g3lhalo/constants.h"

namespace drivers {

struct Tables1D2D {
  std::vector<double> g, p1, p2, w, dwdz, hmf, plin, bh, conc;
};

// Reads the 1-D tables on the redshift grid and the 2-D tables on the
// (z, m) or (z, k) grids, resampled to Nbins as the reference drivers do.
inline void read_tables(int Nbins, double zmin, double zmax, double kmin, double kmax,
                        double mmin, double mmax, const std::string &fn_g,
                        const std::string &fn_p1, const std::string &fn_p2,
                        const std::string &fn_w, const std::string &fn_dwdz,
                        const std::string &fn_hmf, const std::string &fn_plin,
                        const std::string &fn_bh, const std::string &fn_conc, Tables1D2D &t) {
  g3lhalo::Function g(fn_g, 0.0), p1(fn_p1, 0.0), p2(fn_p2, 0.0), w(fn_w, 0.0),
      dwdz(fn_dwdz, 0.0);
  g3lhalo::Function2D hmf(fn_hmf, 0.0), plin(fn_plin, 0.0), bh(fn_bh, 0.0), conc(fn_conc, 0.0);
  if (g.read(Nbins, zmin, zmax, t.g) || p1.read(Nbins, zmin, zmax, t.p1) ||
      p2.read(Nbins, zmin, zmax, t.p2) || w.read(Nbins, zmin, zmax, t.w) ||
      dwdz.read(Nbins, zmin, zmax, t.dwdz) ||
      hmf.read(Nbins, zmin, zmax, mmin, mmax, t.hmf) ||
      plin.read(Nbins, zmin, zmax, kmin, kmax, t.plin) ||
      bh.read(Nbins, zmin, zmax, mmin, mmax, t.bh) ||
      conc.read(Nbins, zmin, zmax, mmin, mmax, t.conc)) {
    std::cerr << "Problem reading in files. Exiting." << std::endl;
    std::exit(1);
  }
}

// Three columns of apertures [arcmin].
inline void read_thetas(const std::string &fn, std::vector<std::array<double, 3>> &arcmin) {
  std::ifstream in(fn);
  double a, b, c;
  while (in >> a >> b >> c) arcmin.push_back({a, b, c});
}

inline std::vector<std::array<double, 3>> to_rad(const std::vector<std::array<double, 3>> &a) {
  std::vector<std::array<double, 3>> r(a.size());
  for (size_t i = 0; i < a.size(); ++i)
    for (int j = 0; j < 3; ++j) r[i][j] = a[i][j] * g3lhalo::arcmin_in_rad;
  return r;
}

// "theta1 theta2 theta3 1h 2h 3h " per line, stream precision 6
// (src/calculateNNMap.cu:148-154).
inline void print_rows(const std::vector<std::array<double, 3>> &arcmin,
                       const std::vector<std::array<double, 3>> &terms) {
  for (size_t i = 0; i < arcmin.size(); ++i)
    std::cout << arcmin[i][0] << " " << arcmin[i][1] << " " << arcmin[i][2] << " "
              << terms[i][0] << " " << terms[i][1] << " " << terms[i][2] << " " << std::endl;
}

}  // namespace drivers
#endif
