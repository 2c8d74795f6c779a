// VUMAT-PD material coupling with the reference's signatures
// (include/VumatPdCorrespondenceMaterial.hpp:11-35 of the reference). The VUMAT runs on the
// GPU as the Johnson-Cook routine (materials/examples/JohnsonCook.f90); a config naming
// another routine is rejected.
#pragma once

#include "PdTypes.hpp"
#include "VumatTypes.hpp"

#include <vector>

namespace vumat {
namespace pd {

enum class KinematicsMode { Corotational, LegacySmallStrain };

struct VumatPdCorrespondenceConfig {
  VumatMaterialConfig material;
  double charLength = 1.0;
  double referenceTemperature = 293.0;
  std::vector<double> initialState;
  KinematicsMode kinematicsMode = KinematicsMode::Corotational;
};

class VumatPdCorrespondenceMaterial {
public:
  explicit VumatPdCorrespondenceMaterial(VumatPdCorrespondenceConfig config);

  void initialize(const std::vector<Tensor3> &initialDeformationGradient,
                  MaterialState &state) const;

  MaterialUpdateResult update(const std::vector<Tensor3> &deformationGradient, MaterialState &state,
                              double dt, double stepTime, double totalTime) const;

private:
  VumatPdCorrespondenceConfig config_;
};

} // namespace pd
} // namespace vumat
