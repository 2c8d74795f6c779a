// Flanagan-Taylor corotational kinematics with the reference's signatures
// (include/CorotationalKinematics.hpp:8-25 of the reference), computed on the GPU.
#pragma once

#include "PdTypes.hpp"

namespace vumat {
namespace pd {

struct CorotationalUpdate {
  Tensor3 velocityGradient{};
  Tensor3 unrotatedRateOfDeformation{};
  Tensor3 relativeSpinIncrement{};
  Tensor3 rotationNP1{};
  Tensor3 leftStretchNP1{};
};

CorotationalUpdate updateCorotationalKinematics(const Tensor3 &deformationGradientN,
                                                const Tensor3 &deformationGradientNP1,
                                                const Tensor3 &leftStretchN,
                                                const Tensor3 &rotationN, double dt);

Tensor3 rotateCorotationalStressToSpatial(const Tensor3 &corotationalStress,
                                          const Tensor3 &rotation);

Tensor3 greenLagrangeStrain(const Tensor3 &deformationGradient);

} // namespace pd
} // namespace vumat
