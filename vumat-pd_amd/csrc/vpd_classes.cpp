// The reference's C++ class API (PdCorrespondenceSkeleton, updateCorotationalKinematics,
// ExplicitTimeIntegrator, VumatPdCorrespondenceMaterial, PdSimpleDamage) implemented over the
// C ABI of include/vpd.h: every computation runs on the GPU; these classes validate as the
// reference does, move data between the reference's std::vector layouts and a device
// context, and turn status codes back into std::runtime_error with the reference's messages.
#include "vpd.h"

#include "CorotationalKinematics.hpp"
#include "ExplicitTimeIntegrator.hpp"
#include "PdCorrespondenceSkeleton.hpp"
#include "PdSimpleDamage.hpp"
#include "VumatPdCorrespondenceMaterial.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

namespace vumat {
namespace pd {
namespace {

void check(int status, const vpd_context *c) {
  if (status != VPD_OK)
    throw std::runtime_error(vpd_last_error(c));
}

// One device context, released on scope exit.
class Ctx {
public:
  explicit Ctx(const vpd_config &cfg) { check(vpd_create(&cfg, &c_), nullptr); }
  ~Ctx() { vpd_destroy(c_); }
  Ctx(const Ctx &) = delete;
  Ctx &operator=(const Ctx &) = delete;
  vpd_context *get() const { return c_; }
  void ok(int status) const { check(status, c_); }

private:
  vpd_context *c_ = nullptr;
};

vpd_config geometryConfig(double horizon) {
  vpd_config cfg{};
  cfg.horizon = horizon;
  return cfg;
}

void validateBlock(const ParticleBlock &block) {
  const size_t n = block.referencePosition.size();
  if (n == 0 || block.displacement.size() != n || block.volume.size() != n ||
      block.neighbors.size() != n)
    throw std::runtime_error("Invalid or incomplete PD particle block.");
}

const double *flat(const std::vector<Vec3> &v) { return v.empty() ? nullptr : v[0].data(); }
const double *flat(const std::vector<Tensor3> &v) { return v.empty() ? nullptr : v[0].data(); }
double *flat(std::vector<Vec3> &v) { return v.empty() ? nullptr : v[0].data(); }
double *flat(std::vector<Tensor3> &v) { return v.empty() ? nullptr : v[0].data(); }

// Uploads positions, volumes and neighbour rows of a block into a context.
void loadBlock(const Ctx &ctx, const ParticleBlock &block) {
  const int64_t n = (int64_t)block.referencePosition.size();
  ctx.ok(vpd_set_particles(ctx.get(), flat(block.referencePosition), block.volume.data(), n));
  std::vector<int32_t> rowptr(n + 1, 0), col;
  for (int64_t i = 0; i < n; ++i) {
    rowptr[i + 1] = rowptr[i] + (int32_t)block.neighbors[i].size();
    for (int j : block.neighbors[i])
      col.push_back(j);
  }
  if (col.empty())
    col.push_back(0);
  ctx.ok(vpd_set_neighbors(ctx.get(), rowptr.data(), col.data(), rowptr[n]));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_DISPLACEMENT, flat(block.displacement)));
}

std::vector<std::vector<int>> gpuNeighbors(const std::vector<Vec3> &positions, double horizon) {
  if (horizon <= 0.0)
    throw std::runtime_error("PD horizon must be positive.");
  const int64_t n = (int64_t)positions.size();
  if (n == 0)
    return {};
  Ctx ctx(geometryConfig(horizon));
  std::vector<double> vol(n, 1.0);
  ctx.ok(vpd_set_particles(ctx.get(), flat(positions), vol.data(), n));
  int64_t nnz = 0;
  ctx.ok(vpd_build_neighbors(ctx.get(), &nnz));
  std::vector<int32_t> rowptr(n + 1), col(std::max<int64_t>(nnz, 1));
  ctx.ok(vpd_get_neighbors(ctx.get(), rowptr.data(), col.data()));
  std::vector<std::vector<int>> out(n);
  for (int64_t i = 0; i < n; ++i)
    out[i].assign(col.begin() + rowptr[i], col.begin() + rowptr[i + 1]);
  return out;
}

} // namespace

// ------------------------------------------------------------------ neighbours

std::vector<std::vector<int>> buildNeighborList(const std::vector<Vec3> &positions,
                                                double horizon) {
  return gpuNeighbors(positions, horizon);
}

std::vector<std::vector<int>> buildNeighborListBruteForce(const std::vector<Vec3> &positions,
                                                          double horizon) {
  return gpuNeighbors(positions, horizon);
}

// ------------------------------------------------------------------ skeleton

PdCorrespondenceSkeleton::PdCorrespondenceSkeleton(double horizon) : horizon_(horizon) {
  if (horizon_ <= 0.0)
    throw std::runtime_error("PD correspondence skeleton requires horizon > 0.");
}

void PdCorrespondenceSkeleton::initialize(const ParticleBlock &block,
                                          CorrespondenceState &state) const {
  validateBlock(block);
  Ctx ctx(geometryConfig(horizon_));
  loadBlock(ctx, block);
  ctx.ok(vpd_initialize_shape(ctx.get()));
  state.shapeTensorInverse.resize(block.referencePosition.size());
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_SHAPE_INVERSE, flat(state.shapeTensorInverse)));
}

std::vector<Tensor3>
PdCorrespondenceSkeleton::computeDeformationGradients(const ParticleBlock &block,
                                                      const CorrespondenceState &state) const {
  validateBlock(block);
  if (state.shapeTensorInverse.size() != block.referencePosition.size())
    throw std::runtime_error("PD correspondence state is not initialized.");
  Ctx ctx(geometryConfig(horizon_));
  loadBlock(ctx, block);
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_SHAPE_INVERSE, flat(state.shapeTensorInverse)));
  ctx.ok(vpd_compute_deformation_gradient(ctx.get()));
  std::vector<Tensor3> f(block.referencePosition.size());
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_DEFGRAD, flat(f)));
  return f;
}

std::vector<Vec3> PdCorrespondenceSkeleton::computeForceDensity(
    const ParticleBlock &block, const CorrespondenceState &state, const std::vector<Tensor3> &f,
    const std::vector<Tensor3> &stress, const std::vector<std::vector<double>> &bondDamage) const {
  const size_t n = block.referencePosition.size();
  if (f.size() != n || stress.size() != n || bondDamage.size() != n)
    throw std::runtime_error("PD force-state inputs do not match particle count.");
  validateBlock(block);
  if (state.shapeTensorInverse.size() != n)
    throw std::runtime_error("PD correspondence state is not initialized.");
  bool anyDamage = false;
  for (size_t i = 0; i < n; ++i) {
    if (bondDamage[i].size() != block.neighbors[i].size())
      throw std::runtime_error("PD bond damage does not match neighbor list.");
    for (double d : bondDamage[i])
      anyDamage = anyDamage || d > 0.0;
  }
  vpd_config cfg = geometryConfig(horizon_);
  if (anyDamage) {
    // Bond flags only enter the force gather; any positive initiation arms them.
    cfg.damage_mode = VPD_DAMAGE_CRITICAL_STRETCH;
    cfg.damage_initiation = 1.0;
  }
  Ctx ctx(cfg);
  loadBlock(ctx, block);
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_SHAPE_INVERSE, flat(state.shapeTensorInverse)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_DEFGRAD, flat(f)));
  if (anyDamage) {
    std::vector<double> flags;
    for (const auto &row : bondDamage)
      flags.insert(flags.end(), row.begin(), row.end());
    ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_BOND_DAMAGE, flags.data()));
  }
  std::vector<Vec3> force(n);
  ctx.ok(vpd_compute_force_density(ctx.get(), flat(stress), flat(force)));
  return force;
}

// ------------------------------------------------------------------ kinematics

CorotationalUpdate updateCorotationalKinematics(const Tensor3 &fN, const Tensor3 &fNP1,
                                                const Tensor3 &vN, const Tensor3 &rN, double dt) {
  CorotationalUpdate out;
  check(vpd_corotational_kinematics(fN.data(), fNP1.data(), vN.data(), rN.data(), dt, 1,
                                    out.velocityGradient.data(),
                                    out.unrotatedRateOfDeformation.data(),
                                    out.relativeSpinIncrement.data(), out.rotationNP1.data(),
                                    out.leftStretchNP1.data()),
        nullptr);
  return out;
}

Tensor3 rotateCorotationalStressToSpatial(const Tensor3 &stress, const Tensor3 &rotation) {
  Tensor3 out{};
  check(vpd_rotate_and_strain(stress.data(), rotation.data(), nullptr, 1, out.data(), nullptr),
        nullptr);
  return out;
}

Tensor3 greenLagrangeStrain(const Tensor3 &f) {
  Tensor3 out{};
  check(vpd_rotate_and_strain(nullptr, nullptr, f.data(), 1, nullptr, out.data()), nullptr);
  return out;
}

// ------------------------------------------------------------------ integrator

ExplicitTimeIntegrator::ExplicitTimeIntegrator(double density) : density_(density) {
  if (density_ <= 0.0)
    throw std::runtime_error("Explicit dynamics requires density > 0.");
}

void ExplicitTimeIntegrator::initialize(const ParticleBlock &block,
                                        ExplicitDynamicState &state) const {
  const size_t n = block.referencePosition.size();
  if (block.displacement.size() != n)
    throw std::runtime_error("Particle displacement size mismatch in explicit dynamics.");
  state.velocity.assign(n, Vec3{0.0, 0.0, 0.0});
  state.acceleration.assign(n, Vec3{0.0, 0.0, 0.0});
}


void ExplicitTimeIntegrator::predictDisplacement(ParticleBlock &block,
                                                 const ExplicitDynamicState &state,
                                                 double dt) const {
  const size_t n = block.displacement.size();
  if (dt <= 0.0 || state.velocity.size() != n || state.acceleration.size() != n)
    throw std::runtime_error("Invalid explicit dynamics predictor input.");
  if (n == 0)
    return;
  vpd_config cfg = geometryConfig(1.0);
  cfg.density = density_;
  Ctx ctx(cfg);
  std::vector<double> x(3 * n, 0.0), vol(n, 1.0);
  ctx.ok(vpd_set_particles(ctx.get(), x.data(), vol.data(), (int64_t)n));
  ctx.ok(vpd_initialize_dynamics(ctx.get()));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_DISPLACEMENT, flat(block.displacement)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_VELOCITY, flat(state.velocity)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_ACCELERATION, flat(state.acceleration)));
  const double none[3] = {0.0, 0.0, 0.0};
  ctx.ok(vpd_predict(ctx.get(), dt, none));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_DISPLACEMENT, flat(block.displacement)));
}

void ExplicitTimeIntegrator::correctVelocity(ExplicitDynamicState &state,
                                             const std::vector<Vec3> &forceDensity,
                                             double dt) const {
  const size_t n = state.velocity.size();
  if (dt <= 0.0 || forceDensity.size() != n || state.acceleration.size() != n)
    throw std::runtime_error("Invalid explicit dynamics corrector input.");
  if (n == 0)
    return;
  vpd_config cfg = geometryConfig(1.0);
  cfg.density = density_;
  Ctx ctx(cfg);
  std::vector<double> x(3 * n, 0.0), vol(n, 1.0);
  ctx.ok(vpd_set_particles(ctx.get(), x.data(), vol.data(), (int64_t)n));
  ctx.ok(vpd_initialize_dynamics(ctx.get()));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_VELOCITY, flat(state.velocity)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_ACCELERATION, flat(state.acceleration)));
  ctx.ok(vpd_correct_velocity(ctx.get(), dt, flat(forceDensity)));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_VELOCITY, flat(state.velocity)));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_ACCELERATION, flat(state.acceleration)));
}

// ------------------------------------------------------------------ material

namespace {
// The device routine a material config names, by its plugin file or cmname; -1 for any
// other. The reference's adapter loads whichever plugin the config names
// (VumatMaterialAdapter.cpp:27-75); the device holds the two example routines
// (materials/examples/JohnsonCook.f90, Elastic.f).
int deviceMaterial(const VumatMaterialConfig &m) {
  std::string plugin = m.pluginPath;
  const size_t slash = plugin.find_last_of("/\\");
  if (slash != std::string::npos)
    plugin = plugin.substr(slash + 1);
  std::string cm = m.cmname;
  std::transform(cm.begin(), cm.end(), cm.begin(), ::toupper);
  if (plugin.rfind("jc_vumat_plugin", 0) == 0 || cm.rfind("JOHNSONCOOK", 0) == 0)
    return VPD_MATERIAL_JOHNSON_COOK;
  if (plugin.rfind("elastic_vumat_plugin", 0) == 0 || cm.rfind("ELASTIC", 0) == 0)
    return VPD_MATERIAL_ELASTIC;
  return -1;
}
} // namespace

VumatPdCorrespondenceMaterial::VumatPdCorrespondenceMaterial(VumatPdCorrespondenceConfig config)
    : config_(std::move(config)) {
  // Order of checks of VumatPdCorrespondenceMaterial.cpp:33-46 and the adapter's
  // constructor (VumatMaterialAdapter.cpp:60-75).
  if (config_.material.ndir != 3 || config_.material.nshr != 3)
    throw std::runtime_error(
        "The minimal adapter currently supports 3D VUMAT only: ndir=3, nshr=3.");
  if (config_.material.nstatev < 0)
    throw std::runtime_error("nstatev must be non-negative.");
  if (config_.material.props.empty())
    throw std::runtime_error("VUMAT props cannot be empty.");
  if (config_.material.tensorConvention != "legacy_vumat_pd")
    throw std::runtime_error("Unsupported tensor convention.");
  if (config_.material.stressFrame != "corotational" && config_.material.stressFrame != "spatial")
    throw std::runtime_error("Unsupported stress frame.");
  const int model = deviceMaterial(config_.material);
  if (model < 0)
    throw std::runtime_error(
        "The GPU material path implements the Johnson-Cook and elastic VUMATs only.");
  if (config_.kinematicsMode != KinematicsMode::Corotational)
    throw std::runtime_error("Only corotational kinematics is implemented by this host.");
  if (config_.material.stressFrame != "corotational")
    throw std::runtime_error("Spatial-stress plugins require the dedicated spatial host.");
  if (config_.charLength <= 0.0)
    throw std::runtime_error("VUMAT-PD material bridge requires charLength > 0.");
  if (!config_.initialState.empty() &&
      static_cast<int>(config_.initialState.size()) != config_.material.nstatev)
    throw std::runtime_error("Initial VUMAT state size does not match nstatev.");
  if (model == VPD_MATERIAL_JOHNSON_COOK &&
      (config_.material.props.size() < 14 || config_.material.nstatev < 6))
    throw std::runtime_error("JohnsonCook VUMAT requires 14 props and 6 states");
  if (model == VPD_MATERIAL_ELASTIC && config_.material.props.size() < 2)
    throw std::runtime_error("The elastic VUMAT requires 2 props (E, nu).");
}

void VumatPdCorrespondenceMaterial::initialize(const std::vector<Tensor3> &f0,
                                               MaterialState &state) const {
  const size_t n = f0.size();
  if (n == 0)
    throw std::runtime_error("VUMAT-PD material bridge requires at least one point.");
  const Tensor3 eye{1, 0, 0, 0, 1, 0, 0, 0, 1};
  state.deformationGradientN = f0;
  state.leftStretchN.assign(n, eye);
  state.rotationN.assign(n, eye);
  state.strainN.resize(n);
  check(vpd_rotate_and_strain(nullptr, nullptr, flat(f0), (int64_t)n, nullptr,
                              flat(state.strainN)),
        nullptr);
  state.rawStressN.assign(n, Tensor3{});
  const std::vector<double> init =
      config_.initialState.empty()
          ? std::vector<double>(static_cast<size_t>(config_.material.nstatev), 0.0)
          : config_.initialState;
  state.vumatStateN.assign(n, init);
  state.temperatureN.assign(n, config_.referenceTemperature);
  state.internalEnergyN.assign(n, 0.0);
  state.inelasticEnergyN.assign(n, 0.0);
}

MaterialUpdateResult VumatPdCorrespondenceMaterial::update(const std::vector<Tensor3> &f,
                                                           MaterialState &state, double dt,
                                                           double stepTime,
                                                           double totalTime) const {
  if (dt <= 0.0)
    throw std::runtime_error("VUMAT-PD material update requires dt > 0.");
  const size_t n = f.size();
  if (state.deformationGradientN.size() != n || state.leftStretchN.size() != n ||
      state.rotationN.size() != n || state.strainN.size() != n || state.rawStressN.size() != n ||
      state.vumatStateN.size() != n || state.temperatureN.size() != n ||
      state.internalEnergyN.size() != n || state.inelasticEnergyN.size() != n)
    throw std::runtime_error("VUMAT material state is not initialized for this point block.");
  const int nst = config_.material.nstatev;
  for (const auto &row : state.vumatStateN)
    if ((int)row.size() != nst)
      throw std::runtime_error("stateOld entry size does not match nstatev.");
  vpd_config cfg = geometryConfig(1.0);
  cfg.material_model = deviceMaterial(config_.material);
  cfg.nprops = (int)config_.material.props.size();
  cfg.props = config_.material.props.data();
  cfg.nstatev = nst;
  cfg.reference_temperature = config_.referenceTemperature;
  cfg.char_length = config_.charLength;
  cfg.material_density = config_.material.density;
  Ctx ctx(cfg);
  std::vector<double> x(3 * n, 0.0), vol(n, 1.0);
  ctx.ok(vpd_set_particles(ctx.get(), x.data(), vol.data(), (int64_t)n));
  // A material-only block: every point is its own (empty) neighbourhood and K⁻¹ = I, so the
  // first Piola map formed beside the update is defined.
  std::vector<int32_t> rowptr(n + 1, 0), col(1, 0);
  ctx.ok(vpd_set_neighbors(ctx.get(), rowptr.data(), col.data(), 0));
  std::vector<Tensor3> eye(n, Tensor3{1, 0, 0, 0, 1, 0, 0, 0, 1});
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_SHAPE_INVERSE, flat(eye)));
  std::vector<double> sv((size_t)n * nst);
  for (size_t i = 0; i < n; ++i)
    std::copy(state.vumatStateN[i].begin(), state.vumatStateN[i].end(), sv.begin() + i * nst);
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_DEFGRAD_N, flat(state.deformationGradientN)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_LEFT_STRETCH_N, flat(state.leftStretchN)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_ROTATION_N, flat(state.rotationN)));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_STRESS_N, flat(state.rawStressN)));
  if (nst > 0)
    ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_STATEV, sv.data()));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_TEMPERATURE, state.temperatureN.data()));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_INTERNAL_ENERGY, state.internalEnergyN.data()));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_INELASTIC_ENERGY, state.inelasticEnergyN.data()));
  ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_DEFGRAD, flat(f)));
  ctx.ok(vpd_update_material(ctx.get(), dt, stepTime, totalTime));

  MaterialUpdateResult out;
  out.deformationGradient = f;
  out.strain.resize(n);
  out.rotation.resize(n);
  out.leftStretch.resize(n);
  out.rawStress.resize(n);
  out.internalEnergy.resize(n);
  out.inelasticEnergy.resize(n);
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_STRESS, flat(out.rawStress)));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_STRAIN, flat(out.strain)));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_ROTATION_N, flat(out.rotation)));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_LEFT_STRETCH_N, flat(out.leftStretch)));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_INTERNAL_ENERGY, out.internalEnergy.data()));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_INELASTIC_ENERGY, out.inelasticEnergy.data()));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_STRESS_N, flat(state.rawStressN)));
  if (nst > 0)
    ctx.ok(vpd_download(ctx.get(), VPD_FIELD_STATEV, sv.data()));
  ctx.ok(vpd_download(ctx.get(), VPD_FIELD_TEMPERATURE, state.temperatureN.data()));
  state.deformationGradientN = f;
  state.leftStretchN = out.leftStretch;
  state.rotationN = out.rotation;
  state.strainN = out.strain;
  for (size_t i = 0; i < n; ++i)
    std::copy(sv.begin() + i * nst, sv.begin() + (i + 1) * nst, state.vumatStateN[i].begin());
  state.internalEnergyN = out.internalEnergy;
  state.inelasticEnergyN = out.inelasticEnergy;
  return out;
}

// ------------------------------------------------------------------ damage

PdSimpleDamage::PdSimpleDamage(SimpleDamageConfig config) : config_(std::move(config)) {
  if (config_.mode != "none" && config_.mode != "critical_stretch")
    throw std::runtime_error("The GPU damage path implements critical_stretch only.");
}

void PdSimpleDamage::initialize(const ParticleBlock &block, PdDamageState &state) const {
  const size_t n = block.referencePosition.size();
  state.bondDamageN.resize(n);
  for (size_t i = 0; i < n; ++i)
    state.bondDamageN[i].assign(block.neighbors[i].size(), 0.0);
  state.damageIndexN.assign(n, 0.0);
  state.pointDamageN.assign(n, SimpleDamageState{});
}

PdDamageResult PdSimpleDamage::update(const ParticleBlock &block,
                                      const std::vector<Tensor3> &strain,
                                      const std::vector<Tensor3> &rawStress,
                                      const std::vector<std::vector<double>> &vumatState,
                                      PdDamageState &state) const {
  const size_t n = block.referencePosition.size();
  if (strain.size() != n || rawStress.size() != n || vumatState.size() != n ||
      state.bondDamageN.size() != n)
    throw std::runtime_error("Simple PD damage inputs do not match particle count.");
  PdDamageResult out;
  out.bondDamage = state.bondDamageN;
  out.effectiveStress = rawStress;
  if (config_.mode == "critical_stretch") {
    if (config_.initiation <= 0.0)
      throw std::runtime_error("critical_stretch requires initiation > 0.");
    validateBlock(block);
    vpd_config cfg = geometryConfig(1.0);
    cfg.damage_mode = VPD_DAMAGE_CRITICAL_STRETCH;
    cfg.damage_initiation = config_.initiation;
    cfg.damage_failure = config_.failure;
    Ctx ctx(cfg);
    loadBlock(ctx, block);
    std::vector<double> flags;
    for (size_t i = 0; i < n; ++i) {
      if (state.bondDamageN[i].size() != block.neighbors[i].size())
        throw std::runtime_error("PD bond damage does not match neighbor list.");
      flags.insert(flags.end(), state.bondDamageN[i].begin(), state.bondDamageN[i].end());
    }
    if (!flags.empty())
      ctx.ok(vpd_upload(ctx.get(), VPD_FIELD_BOND_DAMAGE, flags.data()));
    ctx.ok(vpd_update_damage(ctx.get()));
    if (!flags.empty())
      ctx.ok(vpd_download(ctx.get(), VPD_FIELD_BOND_DAMAGE, flags.data()));
    out.damageIndex.resize(n);
    ctx.ok(vpd_download(ctx.get(), VPD_FIELD_DAMAGE_INDEX, out.damageIndex.data()));
    size_t k = 0;
    for (size_t i = 0; i < n; ++i)
      for (double &d : out.bondDamage[i])
        d = flags[k++];
    for (size_t i = 0; i < n; ++i) {
      state.pointDamageN[i].damage = out.damageIndex[i];
      state.pointDamageN[i].deleted = config_.deletionEnabled && out.damageIndex[i] >= 1.0;
    }
  } else {
    // Mode "none": the point model leaves damage at zero and stress unchanged.
    out.damageIndex.assign(n, 0.0);
    for (size_t i = 0; i < n; ++i)
      state.pointDamageN[i] = SimpleDamageState{};
  }
  state.bondDamageN = out.bondDamage;
  state.damageIndexN = out.damageIndex;
  return out;
}

} // namespace pd
} // namespace vumat
