// Outer damage configuration (the reference's include/SimpleDamageModel.hpp:14-28 fields).
// The GPU path implements the critical_stretch mode; the point modes are not on it.
#pragma once

#include <string>

namespace vumat {

struct SimpleDamageConfig {
  std::string mode = "none";
  double initiation = 0.0;
  double failure = 1.0;
  int eqpsStateIndex = 1;
  bool deletionEnabled = false;
};

struct SimpleDamageState {
  double damage = 0.0;
  bool deleted = false;
};

} // namespace vumat
