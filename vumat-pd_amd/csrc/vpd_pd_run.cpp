// Command line of the GPU case runner, after src/apps/main_pd_case_runner.cpp:15-27:
// vpd_pd_run [case.json], run from the checkout whose relative paths the case names.
#include "vpd.h"

#include <iostream>
#include <string>

int main(int argc, char **argv) {
  if (argc > 1 && (std::string(argv[1]) == "--help" || std::string(argv[1]) == "-h")) {
    std::cout << "Usage: vpd_pd_run [case.json]\n"
              << "Runs a vumat_pd_boundary_case file on the GPU; writes timeseries.csv,\n"
              << "field_snapshots.csv and metadata.json to the case's output.directory.\n"
              << "Units: mm, ms, N, MPa. dt is user supplied.\n";
    return 0;
  }
  if (argc < 2) {
    std::cerr << "VUMAT-PD demo case failed: a vumat_pd_boundary_case JSON path is required\n";
    return 1;
  }
  const std::string self = argv[0];
  const size_t cut = self.find_last_of("\\/");
  const std::string dir = cut == std::string::npos ? "." : self.substr(0, cut);
  return vpd_run_case_json(argv[1], dir.c_str());
}
