// Params, HOD and Cosmology of the C++ API (reference: src/Params.cpp,
// src/HOD.cu, src/Cosmology.cpp:11-87).
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "../../include/g3lhalo/Cosmology.h"
#include "../../include/g3lhalo/HOD.h"
#include "../../include/g3lhalo/Params.h"

namespace {
// "name value" lines; '#' lines and empty lines skipped.
bool read_pairs(const std::string &fn, std::vector<std::string> &names,
                std::vector<double> &values) {
  std::ifstream in(fn.c_str());
  if (in.fail()) return false;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream iss(line);
    std::string name;
    double v = 0;
    iss >> name >> v;
    names.push_back(name);
    values.push_back(v);
  }
  return true;
}
}  // namespace

g3lhalo::Params::Params(const std::string &filename) {
  std::vector<std::string> names;
  std::vector<double> values;
  if (!read_pairs(filename, names, values)) {
    std::cerr << "Params: Could not open " << filename << std::endl;
    std::cerr << "Exiting" << std::endl;
    std::exit(1);
  }
  for (size_t i = 0; i < names.size(); ++i) {
    const std::string &n = names[i];
    if (n == "M_min") mmin = values[i];
    else if (n == "sigma") sigma = values[i];
    else if (n == "M'") mprime = values[i];
    else if (n == "beta") beta = values[i];
    else if (n == "alpha") alpha = values[i];
    else if (n == "f") f = values[i];
    else {
      std::cerr << "Params: Parameter file is not in the right format" << std::endl;
      std::cerr << "Check " << filename << std::endl;
      std::cerr << "Exiting" << std::endl;
      std::exit(1);
    }
  }
}

std::ostream &operator<<(std::ostream &out, g3lhalo::Params const &p) {
  out << "M_min " << p.mmin << std::endl;
  out << "sigma " << p.sigma << std::endl;
  out << "M' " << p.mprime << std::endl;
  out << "beta " << p.beta << std::endl;
  out << "alpha " << p.alpha << std::endl;
  out << "f " << p.f << std::endl;
  return out;
}

double g3lhalo::HOD::Nsat(double m, double *) const {
  return 0.5 * (1 + std::erf(std::log(m / params->mmin) / params->sigma / 1.414213562)) *
         std::pow(m / params->mprime, params->beta);
}

double g3lhalo::HOD::Ncen(double m, double *) const {
  return 0.5 * params->alpha *
         (1 + std::erf(std::log(m / params->mmin) / params->sigma / 1.414213562));
}

double g3lhalo::NsatNsat(double m, HOD *hod1, HOD *hod2, double *, double *, double A,
                         double epsilon, double scale1, double) {
  double ns1 = hod1->Nsat(m);
  if (hod1 == hod2) return (ns1 + scale1 * scale1 - 1) * ns1;
  double ns2 = hod2->Nsat(m);
  double r = ns1 * ns2 + A * std::pow(m, epsilon) * std::sqrt(ns1 * ns2);
  return r < 0 ? 0 : r;
}

g3lhalo::Cosmology::Cosmology(const std::string &fn) {
  std::vector<std::string> names;
  std::vector<double> values;
  if (!read_pairs(fn, names, values)) {
    std::cout << "Cosmology: Could not open " << fn << std::endl;
    return;
  }
  for (size_t i = 0; i < names.size(); ++i) {
    const std::string &n = names[i];
    if (n == "Omega_m") Omega_m = values[i];
    else if (n == "Omega_Lambda") Omega_Lambda = values[i];
    else if (n == "Omega_b") Omega_b = values[i];
    else if (n == "sigma8") sigma8 = values[i];
    else if (n == "H0") H0 = values[i];
    else if (n == "ns") ns = values[i];
    else {
      std::cout << "Cosmology::Parameter file is not in the right format" << std::endl;
      return;
    }
  }
  rho_crit = 3 * H0 * H0 * Mpc_in_km / 8 / g3lhalo::pi / G;
}

double g3lhalo::Cosmology::E(const double &z) const {
  return Omega_m * (1 + z) * (1 + z) * (1 + z) + Omega_Lambda;
}

double g3lhalo::Cosmology::Omega_m_(const double &z) const {
  return Omega_m * (1 + z) * (1 + z) * (1 + z) / E(z);
}
