// Particle-block and history types of the PD increment, with the reference's member names
// and layouts (include/PdTypes.hpp:14-47 of the reference), so callers recompile unchanged.
#pragma once

#include <array>
#include <vector>

namespace vumat {
namespace pd {

using Vec3 = std::array<double, 3>;
using Tensor3 = std::array<double, 9>; // row-major 3x3

struct ParticleBlock {
  std::vector<Vec3> referencePosition;
  std::vector<Vec3> displacement;
  std::vector<double> volume;
  std::vector<std::vector<int>> neighbors; // ascending, i excluded
};

struct CorrespondenceState {
  std::vector<Tensor3> shapeTensorInverse;
};

struct MaterialState {
  std::vector<Tensor3> deformationGradientN;
  std::vector<Tensor3> leftStretchN;
  std::vector<Tensor3> rotationN;
  std::vector<Tensor3> strainN;
  std::vector<Tensor3> rawStressN; // corotational frame
  std::vector<std::vector<double>> vumatStateN;
  std::vector<double> temperatureN;
  std::vector<double> internalEnergyN;
  std::vector<double> inelasticEnergyN;
};

struct MaterialUpdateResult {
  std::vector<Tensor3> deformationGradient;
  std::vector<Tensor3> strain;
  std::vector<Tensor3> rotation;
  std::vector<Tensor3> leftStretch;
  std::vector<Tensor3> rawStress; // spatial Cauchy stress
  std::vector<double> internalEnergy;
  std::vector<double> inelasticEnergy;
};

} // namespace pd
} // namespace vumat
