// Table ingest and command-line helpers of the C++ API (reference:
// src/Function.cpp, src/Function2D.cpp, src/helpers.cpp).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "../../include/g3lhalo/Function.h"
#include "../../include/g3lhalo/Function2D.h"
#include "../../include/g3lhalo/helpers.h"

g3lhalo::Function::Function(std::string filename, const double &default_value) {
  if (filename != "none") {
    std::ifstream input(filename.c_str());
    if (!input.is_open()) {
      std::cerr << "Function: Could not open input file:" << filename << " Exiting. \n";
      std::exit(1);
    }
    double x, y;
    while (input >> x >> y) {
      x_values_.push_back(x);
      y_values_.push_back(y);
    }
    if (x_values_.empty()) {
      std::cerr << "Function: Could not read values. Is file " << filename
                << " empty or contains nans? Exiting." << std::endl;
      std::exit(1);
    }
  }
  default_value_ = default_value;
}

double g3lhalo::Function::at(const double &x) {
  if (x_values_.empty()) return default_value_;
  if (x > x_values_.back() || x < x_values_.front()) return default_value_;
  for (size_t i = 0; i + 1 < x_values_.size(); ++i) {
    if (x_values_[i] <= x && x_values_[i + 1] > x) {
      double a = (y_values_[i + 1] - y_values_[i]) / (x_values_[i + 1] - x_values_[i]);
      return a * (x - x_values_[i]) + y_values_[i];
    }
  }
  return default_value_;  // x equal to the last abscissa
}

double g3lhalo::Function::x_lower(const double &y) {
  for (size_t i = 0; i + 1 < x_values_.size(); ++i)
    if (y_values_[i] <= y && y_values_[i + 1] > y) return x_values_[i];
  return x_values_.back();
}

int g3lhalo::Function::read(const int &Nbins, const double &min, const double &max,
                            std::vector<double> &values) {
  double bin = (max - min) / Nbins;
  for (int i = 0; i < Nbins; ++i) values.push_back(at(min + i * bin));
  if ((int)values.size() != Nbins) {
    std::cerr << "Function::read: Couldn't read function! Exiting." << std::endl;
    std::exit(1);
  }
  return 0;
}

int g3lhalo::Function::readLog(const int &Nbins, const double &min, const double &max,
                               std::vector<double> &values) {
  double bin = std::log(max / min) / Nbins;
  for (int i = 0; i < Nbins; ++i) values.push_back(at(std::exp(std::log(min) + i * bin)));
  if ((int)values.size() != Nbins) {
    std::cerr << "Function::readLog: Couldn't read function! Exiting." << std::endl;
    std::exit(1);
  }
  return 0;
}

g3lhalo::Function2D::Function2D(std::string filename, const double &default_value) {
  if (filename != "none") {
    std::ifstream input(filename.c_str());
    if (!input.is_open()) {
      std::cerr << "Function: Could not open input file:" << filename << " Exiting. \n";
      std::exit(1);
    }
    double x1, x2, y;
    while (input >> x1 >> x2 >> y) {
      x1_values_.push_back(x1);
      x2_values_.push_back(x2);
      y_values_.push_back(y);
    }
    if (input.bad()) {
      std::cerr << "Function2D: Invalid values in " << filename << " Exiting. \n";
      std::exit(1);
    }
  }
  default_value_ = default_value;
  if (!x1_values_.empty()) {
    x1_min = *std::min_element(x1_values_.begin(), x1_values_.end());
    x1_max = *std::max_element(x1_values_.begin(), x1_values_.end());
    x2_min = *std::min_element(x2_values_.begin(), x2_values_.end());
    x2_max = *std::max_element(x2_values_.begin(), x2_values_.end());
  }
  N1 = (int)std::sqrt((double)x1_values_.size());
  N2 = (int)std::sqrt((double)x2_values_.size());
}

double g3lhalo::Function2D::at(const double &x1, const double &x2) {
  if (x1_values_.empty() || x2_values_.empty()) return default_value_;
  int ix1 = (int)std::round((x1 - x1_min) * N1 / (x1_max - x1_min));
  int ix2 = (int)std::round(std::log(x2 / x2_min) * N2 / std::log(x2_max / x2_min));
  if (ix1 > N1 || ix2 > N2) return default_value_;
  if (ix1 < 0 || ix2 < 0) return default_value_;
  return y_values_.at((size_t)ix1 * N2 + ix2);  // throws past the end, as the reference
}

int g3lhalo::Function2D::read(const int &Nbins, const double &min1, const double &max1,
                              const double &min2, const double &max2,
                              std::vector<double> &values) {
  x1_min = min1;
  x1_max = max1;
  x2_min = min2;
  x2_max = max2;
  N1 = Nbins;
  N2 = Nbins;
  double bin1 = (max1 - min1) / Nbins;
  double bin2 = std::log(max2 / min2) / Nbins;
  for (int i = 0; i < Nbins; ++i) {
    double x1 = min1 + i * bin1;
    for (int j = 0; j < Nbins; ++j) values.push_back(at(x1, std::exp(std::log(min2) + j * bin2)));
  }
  if (values.size() != (size_t)Nbins * Nbins) {
    std::cerr << "Function2D::read: Couldn't read function! Exiting." << std::endl;
    std::exit(1);
  }
  return 0;
}

void g3lhalo::checkCmdLine(int argc, int n_params, std::string usage, std::string example) {
  if (argc != n_params + 1) {
    std::cerr << "Wrong number of CMD Line arguments" << std::endl;
    std::cerr << "Expected:" << n_params + 1 << std::endl;
    std::cerr << "Got:" << argc << std::endl;
    std::cerr << "Usage:" << usage << std::endl;
    std::cerr << "Example:" << example << std::endl;
    std::exit(1);
  }
}

void g3lhalo::writeToFile(const std::vector<double> &x, const std::vector<double> &y,
                          const std::string &filename) {
  std::ofstream file(filename);
  for (size_t i = 0; i < x.size(); ++i) file << x.at(i) << " " << y.at(i) << std::endl;
}

void g3lhalo::writeToFile(const std::vector<double> &x1, const std::vector<double> &x2,
                          const std::vector<double> &y, const std::string &filename) {
  std::ofstream file(filename);
  const size_t n = x1.size();
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) file << x1.at(i) << " " << x2.at(j) << " " << y.at(i * n + j)
                                        << std::endl;
}
