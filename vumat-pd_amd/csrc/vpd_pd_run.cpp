// Command line of the GPU case runner, after src/apps/main_pd_case_runner.cpp:15-27:
// vpd_pd_run [case.json], run from the checkout whose relative paths the case names (default
// examples/affine_bar_elastic.json, as the reference's). Given several boundary-case files it
// runs them as one batch on the GPU (vpd_run_cases_json) instead of one process per case.
#include "vpd.h"

#include <iostream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
  if (argc > 1 && (std::string(argv[1]) == "--help" || std::string(argv[1]) == "-h")) {
    std::cout << "Usage: vpd_pd_run [case.json [case.json ...]]\n"
              << "Run from the checkout root; default: examples/affine_bar_elastic.json\n"
              << "Runs a case file on the GPU and writes its outputs to the case's\n"
              << "output.directory. Several vumat_pd_boundary_case files run as one batch.\n"
              << "Regional PD units: mm, ms, N, MPa. dt is user supplied.\n";
    return 0;
  }
  const std::string self = argv[0];
  const size_t cut = self.find_last_of("\\/");
  const std::string dir = cut == std::string::npos ? "." : self.substr(0, cut);
  if (argc > 2) {
    std::vector<const char *> paths(argv + 1, argv + argc);
    return vpd_run_cases_json(paths.data(), (int)paths.size(), dir.c_str());
  }
  const std::string casePath = argc > 1 ? argv[1] : "examples/affine_bar_elastic.json";
  return vpd_run_case_json(casePath.c_str(), dir.c_str());
}
