// Outer PD damage with the reference's signatures (include/PdSimpleDamage.hpp:11-35 of the
// reference). critical_stretch runs on the GPU; mode "none" leaves bonds intact.
#pragma once

#include "PdTypes.hpp"
#include "SimpleDamageModel.hpp"

#include <vector>

namespace vumat {
namespace pd {

struct PdDamageState {
  std::vector<std::vector<double>> bondDamageN;
  std::vector<double> damageIndexN;
  std::vector<SimpleDamageState> pointDamageN;
};

struct PdDamageResult {
  std::vector<std::vector<double>> bondDamage;
  std::vector<double> damageIndex;
  std::vector<Tensor3> effectiveStress;
};

class PdSimpleDamage {
public:
  explicit PdSimpleDamage(SimpleDamageConfig config);
  void initialize(const ParticleBlock &block, PdDamageState &state) const;
  PdDamageResult update(const ParticleBlock &block, const std::vector<Tensor3> &strain,
                        const std::vector<Tensor3> &rawStress,
                        const std::vector<std::vector<double>> &vumatState,
                        PdDamageState &state) const;

private:
  SimpleDamageConfig config_;
};

} // namespace pd
} // namespace vumat
