// VUMAT material configuration (the reference's include/VumatTypes.hpp:8-22 fields).
#pragma once

#include <string>
#include <vector>

namespace vumat {

struct VumatMaterialConfig {
  std::string name;
  std::string cmname;
  std::string pluginPath;
  std::string symbolName = "vumat_adapter_call";
  std::string tensorConvention = "legacy_vumat_pd";
  std::string stressFrame = "corotational";
  int ndir = 3;
  int nshr = 3;
  int nstatev = 0;
  int nfieldv = 1;
  std::vector<double> props;
  double density = 0.0;
};

} // namespace vumat
