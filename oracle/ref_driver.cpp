// ORACLE — test infrastructure only. C entry points over the reference's own C++ classes,
// compiled together with the reference sources (oracle/Makefile, target `ref`) into
// oracle/_ref/libvpd_ref.so. Used by tests to pin the restated oracle and the HIP path
// against the reference itself, and by the benchmark's CPU-baseline leg to time the
// reference's OpenMP increment. No reference source is copied here: this file only
// calls the reference API declared in its headers (include/*.hpp of the reference).
//
// Arrays cross this boundary as the reference holds them: AoS rows of Vec3 / Tensor3.
#include "BoundaryConditions.hpp"
#include "CorotationalKinematics.hpp"
#include "DemoCaseRunner.hpp"
#include "ExplicitTimeIntegrator.hpp"
#include "Geometry.hpp"
#include "PdCorrespondenceSkeleton.hpp"
#include "PdSimpleDamage.hpp"
#include "VumatMaterialAdapter.hpp"
#include "VumatPdCorrespondenceMaterial.hpp"

#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

using vumat::pd::Tensor3;
using vumat::pd::Vec3;

namespace {
thread_local std::string g_error;

std::vector<Vec3> vecs(const double *p, long n) {
  std::vector<Vec3> out(n);
  std::memcpy(out.data(), p, sizeof(double) * 3 * n);
  return out;
}
std::vector<Tensor3> tens(const double *p, long n) {
  std::vector<Tensor3> out(n);
  std::memcpy(out.data(), p, sizeof(double) * 9 * n);
  return out;
}
std::vector<std::vector<int>> rows(const int *rowptr, const int *col, long n) {
  std::vector<std::vector<int>> out(n);
  for (long i = 0; i < n; ++i)
    out[i].assign(col + rowptr[i], col + rowptr[i + 1]);
  return out;
}
long exportRows(const std::vector<std::vector<int>> &r, int *rowptr, int *col, long cap) {
  long nnz = 0;
  rowptr[0] = 0;
  for (size_t i = 0; i < r.size(); ++i) {
    nnz += (long)r[i].size();
    rowptr[i + 1] = (int)nnz;
  }
  if (cap >= nnz)
    for (size_t i = 0; i < r.size(); ++i)
      std::copy(r[i].begin(), r[i].end(), col + rowptr[i]);
  return nnz;
}
vumat::pd::ParticleBlock makeBlock(const double *x, const double *u, const double *vol, long n,
                                   const int *rowptr, const int *col) {
  vumat::pd::ParticleBlock b;
  b.referencePosition = vecs(x, n);
  b.displacement = u ? vecs(u, n) : std::vector<Vec3>(n, Vec3{0, 0, 0});
  b.volume.assign(vol, vol + n);
  b.neighbors = rows(rowptr, col, n);
  return b;
}
vumat::VumatMaterialConfig materialConfig(const char *plugin, int nstatev, int nprops,
                                          const double *props, double density) {
  vumat::VumatMaterialConfig c;
  c.name = "JC";
  c.cmname = "JOHNSONCOOK";
  c.pluginPath = plugin;
  c.nstatev = nstatev;
  c.props.assign(props, props + nprops);
  c.density = density;
  return c;
}

// One reference bar: the objects runMaterial builds (BoundaryValueCaseRunner.cpp:142-190)
// and its step order (:207-229), diagnostics left out.
struct RefCase {
  vumat::demo::BoundaryConditionConfig bc;
  vumat::pd::ParticleBlock block;
  std::unique_ptr<vumat::demo::RegionalConstraintSystem> constraints;
  std::unique_ptr<vumat::pd::PdCorrespondenceSkeleton> skeleton;
  vumat::pd::CorrespondenceState corr;
  std::unique_ptr<vumat::pd::VumatPdCorrespondenceMaterial> bridge;
  vumat::pd::MaterialState mat;
  std::unique_ptr<vumat::pd::PdSimpleDamage> damage;
  vumat::pd::PdDamageState dstate;
  std::unique_ptr<vumat::pd::ExplicitTimeIntegrator> integrator;
  vumat::pd::ExplicitDynamicState dyn;
  vumat::pd::MaterialUpdateResult last;
  vumat::pd::PdDamageResult lastDamage;
  std::vector<Vec3> lastForce;
  double dt = 0.0;
  int step = 0;
};
} // namespace

extern "C" {

const char *ref_last_error() {
  return g_error.c_str();
}

long ref_build_neighbors(const double *x, long n, double horizon, int brute, int *rowptr,
                         int *col, long cap) {
  try {
    auto p = vecs(x, n);
    return exportRows(brute ? vumat::pd::buildNeighborListBruteForce(p, horizon)
                            : vumat::pd::buildNeighborList(p, horizon),
                      rowptr, col, cap);
  } catch (const std::exception &e) {
    g_error = e.what();
    return -1;
  }
}

int ref_shape_inverse(const double *x, const double *vol, long n, const int *rowptr,
                      const int *col, double horizon, double *kinv) {
  try {
    auto b = makeBlock(x, nullptr, vol, n, rowptr, col);
    vumat::pd::PdCorrespondenceSkeleton s(horizon);
    vumat::pd::CorrespondenceState st;
    s.initialize(b, st);
    std::memcpy(kinv, st.shapeTensorInverse.data(), sizeof(double) * 9 * n);
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

int ref_deformation_gradient(const double *x, const double *u, const double *vol, long n,
                             const int *rowptr, const int *col, double horizon,
                             const double *kinv, double *f) {
  try {
    auto b = makeBlock(x, u, vol, n, rowptr, col);
    vumat::pd::PdCorrespondenceSkeleton s(horizon);
    vumat::pd::CorrespondenceState st;
    st.shapeTensorInverse = tens(kinv, n);
    auto out = s.computeDeformationGradients(b, st);
    std::memcpy(f, out.data(), sizeof(double) * 9 * n);
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

int ref_force_density(const double *x, const double *vol, long n, const int *rowptr,
                      const int *col, double horizon, const double *kinv, const double *f,
                      const double *sigma, const double *bondDamage, double *force) {
  try {
    auto b = makeBlock(x, nullptr, vol, n, rowptr, col);
    vumat::pd::PdCorrespondenceSkeleton s(horizon);
    vumat::pd::CorrespondenceState st;
    st.shapeTensorInverse = tens(kinv, n);
    std::vector<std::vector<double>> dmg(n);
    for (long i = 0; i < n; ++i)
      dmg[i].assign(bondDamage ? bondDamage + rowptr[i] : nullptr,
                    bondDamage ? bondDamage + rowptr[i + 1] : nullptr);
    if (!bondDamage)
      for (long i = 0; i < n; ++i)
        dmg[i].assign(rowptr[i + 1] - rowptr[i], 0.0);
    auto out = s.computeForceDensity(b, st, tens(f, n), tens(sigma, n), dmg);
    std::memcpy(force, out.data(), sizeof(double) * 3 * n);
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

int ref_corotational(const double *fN, const double *fNP1, const double *vN, const double *rN,
                     double dt, long n, double *L, double *Dbar, double *relSpin, double *R1,
                     double *V1) {
  try {
    for (long i = 0; i < n; ++i) {
      Tensor3 a, b, c, d;
      std::memcpy(a.data(), fN + 9 * i, 72);
      std::memcpy(b.data(), fNP1 + 9 * i, 72);
      std::memcpy(c.data(), vN + 9 * i, 72);
      std::memcpy(d.data(), rN + 9 * i, 72);
      auto k = vumat::pd::updateCorotationalKinematics(a, b, c, d, dt);
      std::memcpy(L + 9 * i, k.velocityGradient.data(), 72);
      std::memcpy(Dbar + 9 * i, k.unrotatedRateOfDeformation.data(), 72);
      std::memcpy(relSpin + 9 * i, k.relativeSpinIncrement.data(), 72);
      std::memcpy(R1 + 9 * i, k.rotationNP1.data(), 72);
      std::memcpy(V1 + 9 * i, k.leftStretchNP1.data(), 72);
    }
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

// VumatMaterialAdapter::updateBlock on one block through the named plugin.
int ref_vumat_block(const char *plugin, int nstatev, int nprops, const double *props,
                    double density, long n, double dt, const double *strainRateFull,
                    const double *stressFullN, const double *stateOld, const double *temp,
                    const double *eint, const double *einel, double *stressFullNew,
                    double *stateNew, double *tempNew, double *eintNew, double *einelNew) {
  try {
    vumat::VumatMaterialAdapter adapter(materialConfig(plugin, nstatev, nprops, props, density));
    vumat::VumatBlockInput in;
    in.dt = dt;
    in.charLength.assign(n, 1.0);
    in.temperature.assign(temp, temp + n);
    in.strainRateFull = tens(strainRateFull, n);
    in.strainFullN.assign(n, Tensor3{});
    in.stressFullN = tens(stressFullN, n);
    in.defgradOld.assign(n, Tensor3{1, 0, 0, 0, 1, 0, 0, 0, 1});
    in.defgradNew = in.defgradOld;
    in.stateOld.resize(n);
    for (long i = 0; i < n; ++i)
      in.stateOld[i].assign(stateOld + (size_t)nstatev * i, stateOld + (size_t)nstatev * (i + 1));
    in.internalEnergyOld.assign(eint, eint + n);
    in.inelasticEnergyOld.assign(einel, einel + n);
    auto r = adapter.updateBlock(in);
    for (long i = 0; i < n; ++i) {
      std::memcpy(stressFullNew + 9 * i, r.stressFullNew[i].data(), 72);
      std::memcpy(stateNew + (size_t)nstatev * i, r.stateNew[i].data(), sizeof(double) * nstatev);
      tempNew[i] = r.temperatureNew[i];
      eintNew[i] = r.internalEnergyNew[i];
      einelNew[i] = r.inelasticEnergyNew[i];
    }
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

// A reference bar case: makeParticlePatch, RegionalConstraintSystem, skeleton, material
// bridge, damage model and integrator built as in BoundaryValueCaseRunner.cpp:142-190.
// component 0 = x (left/right), 1 = y (bottom/top). damageMode 0 none, 1 critical_stretch.
void *ref_case_create(int nx, int ny, int nz, double spacing, double horizon, int component,
                      int regionLayers, double amplitude, double riseTime, int planeStrain,
                      const char *plugin, int nstatev, int nprops, const double *props,
                      double materialDensity, double tref, const double *initialState,
                      double density, int damageMode, double initiation, double failure,
                      double dt) {
  try {
    auto c = std::make_unique<RefCase>();
    c->bc.type = "regional_smooth_displacement";
    c->bc.mode = "boundary_value";
    c->bc.component = component == 0 ? "x" : "y";
    c->bc.supportRegion = component == 0 ? "left" : "bottom";
    c->bc.loadRegion = component == 0 ? "right" : "top";
    c->bc.regionLayers = regionLayers;
    c->bc.amplitude = amplitude;
    c->bc.riseTime = riseTime;
    c->bc.planeStrain = planeStrain != 0;
    c->bc.rigidBodySuppression = true;
    c->block = vumat::demo::makeParticlePatch(nx, ny, nz, spacing, horizon);
    c->constraints =
        std::make_unique<vumat::demo::RegionalConstraintSystem>(c->block, c->bc, spacing);
    c->skeleton = std::make_unique<vumat::pd::PdCorrespondenceSkeleton>(horizon);
    c->skeleton->initialize(c->block, c->corr);
    vumat::pd::VumatPdCorrespondenceConfig mc;
    mc.material = materialConfig(plugin, nstatev, nprops, props, materialDensity);
    mc.charLength = spacing;
    mc.referenceTemperature = tref;
    mc.initialState = initialState ? std::vector<double>(initialState, initialState + nstatev)
                                   : std::vector<double>(nstatev, 0.0);
    c->bridge = std::make_unique<vumat::pd::VumatPdCorrespondenceMaterial>(mc);
    c->bridge->initialize(c->skeleton->computeDeformationGradients(c->block, c->corr), c->mat);
    vumat::SimpleDamageConfig dc;
    if (damageMode == 1) {
      dc.mode = "critical_stretch";
      dc.initiation = initiation;
      dc.failure = failure;
    }
    c->damage = std::make_unique<vumat::pd::PdSimpleDamage>(dc);
    c->damage->initialize(c->block, c->dstate);
    c->integrator = std::make_unique<vumat::pd::ExplicitTimeIntegrator>(density);
    c->integrator->initialize(c->block, c->dyn);
    c->constraints->apply(c->block, c->dyn, 0.0);
    c->dt = dt;
    return c.release();
  } catch (const std::exception &e) {
    g_error = e.what();
    return nullptr;
  }
}

void ref_case_destroy(void *h) {
  delete static_cast<RefCase *>(h);
}

long ref_case_size(void *h, long *nnz) {
  auto *c = static_cast<RefCase *>(h);
  long b = 0;
  for (auto &r : c->block.neighbors)
    b += (long)r.size();
  if (nnz)
    *nnz = b;
  return (long)c->block.referencePosition.size();
}

// Runs `steps` increments in the order of BoundaryValueCaseRunner.cpp:207-229.
int ref_case_step(void *h, int steps) {
  auto *c = static_cast<RefCase *>(h);
  try {
    for (int s = 0; s < steps; ++s) {
      const int step = ++c->step;
      const double time = static_cast<double>(step) * c->dt;
      c->integrator->predictDisplacement(c->block, c->dyn, c->dt);
      c->constraints->apply(c->block, c->dyn, time);
      const auto f = c->skeleton->computeDeformationGradients(c->block, c->corr);
      c->last = c->bridge->update(f, c->mat, c->dt, (step - 1) * c->dt, (step - 1) * c->dt);
      c->lastDamage = c->damage->update(c->block, c->last.strain, c->last.rawStress,
                                        c->mat.vumatStateN, c->dstate);
      c->lastForce = c->skeleton->computeForceDensity(c->block, c->corr, c->last.deformationGradient,
                                                      c->lastDamage.effectiveStress,
                                                      c->lastDamage.bondDamage);
      c->integrator->correctVelocity(c->dyn, c->lastForce, c->dt);
      c->constraints->apply(c->block, c->dyn, time);
    }
    return 0;
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

// Field export. id: 0 x, 1 volume, 2 u, 3 v, 4 a, 5 F_n, 6 V_n, 7 R_n, 8 σ̄_n (corotational),
// 9 statev (n×nstatev), 10 temperature, 11 internal energy, 12 inelastic energy,
// 13 σ (spatial, last step), 14 strain (last step), 15 force density (last step),
// 16 bond damage (per CSR entry), 17 damage index (last step), 18 K⁻¹, 19 rowptr (int),
// 20 col (int).
int ref_case_get(void *h, int id, void *out) {
  auto *c = static_cast<RefCase *>(h);
  const long n = (long)c->block.referencePosition.size();
  double *d = static_cast<double *>(out);
  auto put3 = [&](const std::vector<Vec3> &v) { std::memcpy(d, v.data(), sizeof(double) * 3 * n); };
  auto put9 = [&](const std::vector<Tensor3> &v) {
    std::memcpy(d, v.data(), sizeof(double) * 9 * n);
  };
  switch (id) {
  case 0:
    put3(c->block.referencePosition);
    return 0;
  case 1:
    std::memcpy(d, c->block.volume.data(), sizeof(double) * n);
    return 0;
  case 2:
    put3(c->block.displacement);
    return 0;
  case 3:
    put3(c->dyn.velocity);
    return 0;
  case 4:
    put3(c->dyn.acceleration);
    return 0;
  case 5:
    put9(c->mat.deformationGradientN);
    return 0;
  case 6:
    put9(c->mat.leftStretchN);
    return 0;
  case 7:
    put9(c->mat.rotationN);
    return 0;
  case 8:
    put9(c->mat.rawStressN);
    return 0;
  case 9: {
    size_t k = 0;
    for (auto &r : c->mat.vumatStateN)
      for (double v : r)
        d[k++] = v;
    return 0;
  }
  case 10:
    std::memcpy(d, c->mat.temperatureN.data(), sizeof(double) * n);
    return 0;
  case 11:
    std::memcpy(d, c->mat.internalEnergyN.data(), sizeof(double) * n);
    return 0;
  case 12:
    std::memcpy(d, c->mat.inelasticEnergyN.data(), sizeof(double) * n);
    return 0;
  case 13:
    if (c->last.rawStress.empty())
      return 1;
    put9(c->last.rawStress);
    return 0;
  case 14:
    if (c->last.strain.empty())
      return 1;
    put9(c->last.strain);
    return 0;
  case 15:
    if (c->lastForce.empty())
      return 1;
    put3(c->lastForce);
    return 0;
  case 16: {
    size_t k = 0;
    for (auto &r : c->dstate.bondDamageN)
      for (double v : r)
        d[k++] = v;
    return 0;
  }
  case 17:
    if (c->lastDamage.damageIndex.empty())
      return 1;
    std::memcpy(d, c->lastDamage.damageIndex.data(), sizeof(double) * n);
    return 0;
  case 18:
    put9(c->corr.shapeTensorInverse);
    return 0;
  case 19: {
    int *r = static_cast<int *>(out);
    long nnz = 0;
    r[0] = 0;
    for (long i = 0; i < n; ++i) {
      nnz += (long)c->block.neighbors[i].size();
      r[i + 1] = (int)nnz;
    }
    return 0;
  }
  case 20: {
    int *r = static_cast<int *>(out);
    size_t k = 0;
    for (auto &row : c->block.neighbors)
      for (int j : row)
        r[k++] = j;
    return 0;
  }
  default:
    return 1;
  }
}

// Field import, the inverse of ref_case_get for the state a step reads: 2 u, 3 v, 4 a,
// 5 F_n (strainN follows as greenLagrangeStrain(F_n), CorotationalKinematics.cpp:141-149, as
// VumatPdCorrespondenceMaterial::update leaves it), 6 V_n, 7 R_n, 8 σ̄_n, 9 statev,
// 10 temperature, 11 internal energy, 12 inelastic energy, 16 bond damage. Lets a sample of
// the reference start from a state reached elsewhere (e.g. a plastic phase).
int ref_case_set(void *h, int id, const void *in) {
  auto *c = static_cast<RefCase *>(h);
  const long n = (long)c->block.referencePosition.size();
  const double *d = static_cast<const double *>(in);
  switch (id) {
  case 2:
    c->block.displacement = vecs(d, n);
    return 0;
  case 3:
    c->dyn.velocity = vecs(d, n);
    return 0;
  case 4:
    c->dyn.acceleration = vecs(d, n);
    return 0;
  case 5:
    c->mat.deformationGradientN = tens(d, n);
    for (long i = 0; i < n; ++i)
      c->mat.strainN[i] = vumat::pd::greenLagrangeStrain(c->mat.deformationGradientN[i]);
    return 0;
  case 6:
    c->mat.leftStretchN = tens(d, n);
    return 0;
  case 7:
    c->mat.rotationN = tens(d, n);
    return 0;
  case 8:
    c->mat.rawStressN = tens(d, n);
    return 0;
  case 9: {
    size_t k = 0;
    for (auto &r : c->mat.vumatStateN)
      for (double &v : r)
        v = d[k++];
    return 0;
  }
  case 10:
    std::memcpy(c->mat.temperatureN.data(), d, sizeof(double) * n);
    return 0;
  case 11:
    std::memcpy(c->mat.internalEnergyN.data(), d, sizeof(double) * n);
    return 0;
  case 12:
    std::memcpy(c->mat.inelasticEnergyN.data(), d, sizeof(double) * n);
    return 0;
  case 16: {
    size_t k = 0;
    for (auto &r : c->dstate.bondDamageN)
      for (double &v : r)
        v = d[k++];
    return 0;
  }
  default:
    return 1;
  }
}

// Sets the number of completed steps (the next ref_case_step runs step + 1).
void ref_case_set_step(void *h, int step) {
  static_cast<RefCase *>(h)->step = step;
}

// The reference CLI path: runDemoCaseFromJson (DemoCaseRunner.cpp:435-492).
int ref_run_case_json(const char *casePath, const char *pluginDirectory) {
  try {
    return vumat::demo::runDemoCaseFromJson(casePath, pluginDirectory);
  } catch (const std::exception &e) {
    g_error = e.what();
    return 1;
  }
}

} // extern "C"
