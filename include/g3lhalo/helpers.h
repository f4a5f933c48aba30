// helpers.h -- command-line check and table writers (src/helpers.h).
#ifndef G3LHALO_HELPERS_H
#define G3LHALO_HELPERS_H

#include <string>
#include <vector>

namespace g3lhalo {
// Exits(1) with usage and example when argc != n_params + 1 (src/helpers.cpp:6-17).
void checkCmdLine(int argc, int n_params, std::string usage, std::string example);
void writeToFile(const std::vector<double> &x, const std::vector<double> &y,
                 const std::string &filename);
void writeToFile(const std::vector<double> &x1, const std::vector<double> &x2,
                 const std::vector<double> &y, const std::string &filename);
}  // namespace g3lhalo
#endif
