// The reference's unit tests, restated against this build's C++ class API (every call runs
// on the GPU through libvpd.so). Each check names the reference test it follows.
#include "CorotationalKinematics.hpp"
#include "ExplicitTimeIntegrator.hpp"
#include "PdCorrespondenceSkeleton.hpp"
#include "PdSimpleDamage.hpp"
#include "VumatPdCorrespondenceMaterial.hpp"
#include "vpd.h"

#include <cmath>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

using namespace vumat;
using namespace vumat::pd;

namespace {
int failures = 0;
void require(bool ok, const std::string &what) {
  if (!ok) {
    std::printf("FAIL %s\n", what.c_str());
    ++failures;
  }
}
template <class F> bool throws(F f, const std::string &needle) {
  try {
    f();
  } catch (const std::runtime_error &e) {
    return std::string(e.what()).find(needle) != std::string::npos;
  }
  return false;
}
double maxDiff(const Tensor3 &a, const Tensor3 &b) {
  double m = 0;
  for (int i = 0; i < 9; ++i)
    m = std::max(m, std::abs(a[i] - b[i]));
  return m;
}
Tensor3 eye() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }

ParticleBlock patch(int nx, int ny, int nz, double h, double horizon) {
  ParticleBlock b;
  const size_t n = (size_t)nx * ny * nz;
  b.referencePosition.resize(n);
  b.volume.resize(n);
  if (vpd_make_patch(nx, ny, nz, h, b.referencePosition[0].data(), b.volume.data()))
    throw std::runtime_error("patch");
  b.displacement.assign(n, Vec3{0, 0, 0});
  b.neighbors = buildNeighborList(b.referencePosition, horizon);
  return b;
}

// tests/neighbor_lattice_regression.cpp:5-36
void neighborLattice() {
  for (double h : {2.0 / 3.0, 1.0, 2.0})
    for (double shift : {0.0, -123.456, 10000.125}) {
      const int n = 9;
      std::vector<Vec3> x;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
          for (int k = 0; k < n; ++k)
            x.push_back({shift + i * h, shift + j * h, shift + k * h});
      auto fast = buildNeighborList(x, 3 * h);
      auto slow = buildNeighborListBruteForce(x, 3 * h);
      bool ok = true;
      for (int i = 0; i < n && ok; ++i)
        for (int j = 0; j < n && ok; ++j)
          for (int k = 0; k < n && ok; ++k) {
            std::vector<int> expected;
            for (int a = 0; a < n; ++a)
              for (int b = 0; b < n; ++b)
                for (int c = 0; c < n; ++c) {
                  int d = (a - i) * (a - i) + (b - j) * (b - j) + (c - k) * (c - k);
                  if (d > 0 && d <= 9)
                    expected.push_back((a * n + b) * n + c);
                }
            const int p = (i * n + j) * n + k;
            ok = fast[p] == expected && slow[p] == expected;
          }
      require(ok, "neighbor lattice h=" + std::to_string(h) + " shift=" + std::to_string(shift));
    }
}

// tests/correspondence_sampling_test.cpp:7-46
void correspondenceSampling() {
  auto block = patch(9, 9, 9, 1.0, 2.0);
  PdCorrespondenceSkeleton pd(2.0);
  CorrespondenceState state;
  pd.initialize(block, state);
  double volume = 0;
  for (double v : block.volume)
    volume += v;
  require(std::abs(volume - 512.0) <= 1e-12, "patch volume 512");
  for (size_t i = 0; i < block.volume.size(); ++i) {
    const auto &x = block.referencePosition[i];
    block.displacement[i] = {0.02 * x[0] + 0.03 * x[1], -0.01 * x[1], 0.04 * x[2]};
  }
  const Tensor3 expected{1.02, 0.03, 0, 0, 0.99, 0, 0, 0, 1.04};
  auto f = pd.computeDeformationGradients(block, state);
  double err = 0;
  for (const auto &t : f)
    err = std::max(err, maxDiff(t, expected));
  require(err <= 1e-12, "affine completeness");
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j)
      for (int k = 0; k < 9; ++k)
        block.displacement[(i * 9 + j) * 9 + k] = {((i + j + k) % 2 ? 1.0 : -1.0) * 0.01, 0, 0};
  f = pd.computeDeformationGradients(block, state);
  err = 0;
  for (int i = 3; i <= 5; ++i)
    for (int j = 3; j <= 5; ++j)
      for (int k = 3; k <= 5; ++k)
        err = std::max(err, maxDiff(f[(i * 9 + j) * 9 + k], eye()));
  require(err <= 1e-12, "checkerboard invisible at interior nodes");
  // Net internal force of a self-equilibrated state (BoundaryValueCaseRunner.cpp:251-252).
  std::vector<Tensor3> sigma(block.volume.size(), Tensor3{100, 5, 3, 5, -20, 7, 3, 7, 40});
  std::vector<std::vector<double>> dmg(block.volume.size());
  for (size_t i = 0; i < dmg.size(); ++i)
    dmg[i].assign(block.neighbors[i].size(), 0.0);
  auto force = pd.computeForceDensity(block, state, f, sigma, dmg);
  double net[3] = {0, 0, 0}, mag = 0;
  for (size_t i = 0; i < force.size(); ++i)
    for (int d = 0; d < 3; ++d) {
      net[d] += force[i][d] * block.volume[i];
      mag += std::abs(force[i][d] * block.volume[i]);
    }
  require(std::abs(net[0]) + std::abs(net[1]) + std::abs(net[2]) <= 1e-9 * mag, "force balance");
  require(throws([&] { PdCorrespondenceSkeleton bad(0.0); }, "horizon > 0"), "horizon check");
}

// tests/corotational_kinematics_test.cpp:28-137
void corotational() {
  const double angle = 1.0e-7;
  const Tensor3 fr = {std::cos(angle), -std::sin(angle), 0, std::sin(angle), std::cos(angle), 0,
                      0, 0, 1};
  auto u = updateCorotationalKinematics(eye(), fr, eye(), eye(), 1.0);
  require(maxDiff(u.unrotatedRateOfDeformation, Tensor3{}) <= 1e-13, "rigid rotation D=0");
  require(maxDiff(u.leftStretchNP1, eye()) <= 1e-13, "rigid rotation V=I");
  u = updateCorotationalKinematics(eye(), {1.01, 0, 0, 0, 1, 0, 0, 0, 1}, eye(), eye(), 0.01);
  require(std::abs(u.unrotatedRateOfDeformation[0] - 1.0) <= 1e-12, "extension D");
  require(std::abs(u.leftStretchNP1[0] - 1.01) <= 1e-12, "extension V");
  require(std::abs(greenLagrangeStrain({1.01, 0, 0, 0, 1, 0, 0, 0, 1})[0] - 0.01005) <= 1e-12,
          "Green-Lagrange");
  const auto s = rotateCorotationalStressToSpatial({10, 0, 0, 0, 2, 0, 0, 0, 1},
                                                   {0, -1, 0, 1, 0, 0, 0, 0, 1});
  require(std::abs(s[0] - 2.0) <= 1e-12 && std::abs(s[4] - 10.0) <= 1e-12, "stress rotation");
  for (int axis = 0; axis < 3; ++axis)
    for (double a : {1.0e-7, 0.01, 0.1, 0.7, -0.4}) {
      const int p = (axis + 1) % 3, q = (axis + 2) % 3;
      Tensor3 f = eye(), e = eye();
      f[3 * p + q] = -a;
      f[3 * q + p] = a;
      e[3 * p + p] = e[3 * q + q] = std::cos(a);
      e[3 * p + q] = -std::sin(a);
      e[3 * q + p] = std::sin(a);
      const auto k = updateCorotationalKinematics(eye(), f, eye(), eye(), 1.0);
      require(maxDiff(k.rotationNP1, e) <= 2.0e-14, "finite spin exponential");
    }
  Tensor3 rot = eye(), prev = eye();
  for (int step = 1; step <= 1000; ++step) {
    const double g = 0.01 * step;
    const Tensor3 f = {std::cos(g), -std::sin(g), 0, std::sin(g), std::cos(g), 0, 0, 0, 1};
    rot = updateCorotationalKinematics(prev, f, eye(), rot, 1.0).rotationNP1;
    prev = f;
  }
  const double r = 1000 * std::sin(0.01);
  require(maxDiff(rot, {std::cos(r), -std::sin(r), 0, std::sin(r), std::cos(r), 0, 0, 0, 1}) <=
              2e-12,
          "accumulated rotation");
  require(throws([&] { updateCorotationalKinematics(eye(), eye(), eye(), eye(), 0.0); },
                 "dt > 0"),
          "dt check");
}

// tests/outer_damage_test.cpp:61-91 (critical stretch)
void outerDamage() {
  ParticleBlock block;
  block.referencePosition = {{0.0, 0.0, 0.0}, {1.0, 0.0, 0.0}};
  block.displacement = {{0.0, 0.0, 0.0}, {0.10, 0.0, 0.0}};
  block.volume = {1.0, 2.0};
  block.neighbors = {{1}, {0}};
  std::vector<Tensor3> strains(2), stresses(2);
  stresses[0].fill(20.0);
  stresses[1].fill(30.0);
  std::vector<std::vector<double>> states(2);
  SimpleDamageConfig config;
  config.mode = "critical_stretch";
  config.initiation = 0.05;
  config.failure = 0.0;
  PdSimpleDamage bonds(config);
  PdDamageState st;
  bonds.initialize(block, st);
  const auto broken = bonds.update(block, strains, stresses, states, st);
  require(broken.bondDamage[0][0] == 1.0 && broken.bondDamage[1][0] == 1.0 &&
              broken.damageIndex[0] == 1.0 && broken.damageIndex[1] == 1.0,
          "critical stretch breaks both directed bonds");
  require(broken.effectiveStress == stresses, "damage leaves stress unchanged");
  block.displacement[1] = {0.0, 0.0, 0.0};
  require(bonds.update(block, strains, stresses, states, st).bondDamage == broken.bondDamage,
          "no healing");
  config.failure = 0.02;
  PdSimpleDamage compression(config);
  compression.initialize(block, st);
  block.displacement[1] = {-0.03, 0.0, 0.0};
  require(compression.update(block, strains, stresses, states, st).bondDamage[0][0] == 1.0,
          "compression threshold");
}

// ExplicitTimeIntegrator.cpp:23-52 closed form.
void integrator() {
  ExplicitTimeIntegrator it(2.0);
  ParticleBlock b;
  b.referencePosition.assign(3, Vec3{0, 0, 0});
  b.displacement.assign(3, Vec3{0, 0, 0});
  ExplicitDynamicState st;
  it.initialize(b, st);
  st.velocity.assign(3, Vec3{2, 2, 2});
  st.acceleration.assign(3, Vec3{3, 3, 3});
  it.predictDisplacement(b, st, 0.5);
  require(b.displacement[1][2] == 0.5 * 2.0 + 0.5 * 0.25 * 3.0, "predictor");
  it.correctVelocity(st, std::vector<Vec3>(3, Vec3{6, 6, 6}), 0.5);
  require(st.acceleration[2][0] == 3.0 && st.velocity[0][1] == 2.0 + 0.25 * 6.0, "corrector");
  require(throws([&] { ExplicitTimeIntegrator bad(0.0); }, "density > 0"), "density check");
}

// tests/jc_thermal_units_test.cpp:7-42 and the host-level part of
// tests/interface_contract_test.cpp:112-136 through VumatPdCorrespondenceMaterial.
void material() {
  auto jc = [](double workToHeat) {
    VumatPdCorrespondenceConfig c;
    c.material.cmname = "JOHNSONCOOK";
    c.material.pluginPath = "jc_vumat_plugin.so";
    c.material.nstatev = 6;
    c.material.props = {210000.0, 0.30, 350.0, 500.0, 0.025, 0.36, 1.0,
                        1.0,      293.0, 1793.0, 7800.0, 0.9, 500.0, workToHeat};
    c.material.density = 7800;
    return c;
  };
  double dT[2];
  Tensor3 s[2];
  for (int w = 0; w < 2; ++w) {
    VumatPdCorrespondenceMaterial m(jc(w ? 1e6 : 1.0));
    MaterialState st;
    m.initialize({eye()}, st);
    m.update({eye()}, st, 0.000025, 0, 0); // first call: init branch
    Tensor3 f = eye();
    f[0] = 1.01; // 0.01 strain increment at the reference temperature
    auto out = m.update({f}, st, 0.000025, 0.000025, 0.000025);
    dT[w] = st.temperatureN[0] - 293.0;
    s[w] = out.rawStress[0];
  }
  require(maxDiff(s[0], s[1]) <= 1e-10, "first plastic step stress independent of workToHeat");
  require(dT[0] > 0 && std::abs(dT[1] / (1e6 * dT[0]) - 1) <= 1e-6, "heat-unit ratio");
  VumatPdCorrespondenceMaterial m(jc(1.0));
  MaterialState st;
  std::vector<Tensor3> grads(257, eye());
  m.initialize(grads, st);
  for (int step = 1; step <= 2; ++step) {
    auto out = m.update(grads, st, 0.125, step * 0.125, step * 0.125);
    require(out.rawStress.size() == 257 && st.vumatStateN.size() == 257, "no padding leak");
    require(st.vumatStateN[256][0] == 1.0, "every point initialised");
  }
  auto bad = jc(1.0);
  bad.kinematicsMode = KinematicsMode::LegacySmallStrain;
  require(throws([&] { VumatPdCorrespondenceMaterial x(bad); }, "corotational"),
          "legacy kinematics rejected");
  bad = jc(1.0);
  bad.material.stressFrame = "spatial";
  require(throws([&] { VumatPdCorrespondenceMaterial x(bad); }, "spatial"),
          "spatial stress frame rejected");
  bad = jc(1.0);
  bad.material.cmname = "PROBE";
  bad.material.pluginPath = "contract_probe_plugin.so";
  require(throws([&] { VumatPdCorrespondenceMaterial x(bad); }, "Johnson-Cook and elastic"),
          "a routine the device does not hold is rejected");
  bad = jc(1.0);
  bad.material.cmname = "ELASTIC";
  bad.material.pluginPath = "elastic_vumat_plugin.so";
  bad.material.props = {210000.0};
  require(throws([&] { VumatPdCorrespondenceMaterial x(bad); }, "2 props"),
          "elastic routine without nu rejected");
}

// The elastic routine (materials/examples/Elastic.f) through the same class, as the
// reference's adapter runs any plugin its config names (VumatMaterialAdapter.cpp:27-75).
// Hooke's law on the Flanagan-Taylor increment: the corotational stress grows by
// λ tr(dε) I + 2μ dε with dε = D̄·dt, and the spatial stress is R σ̄ Rᵀ; the state,
// the temperature and the energies are carried unchanged. A general (non-symmetric) F makes
// every shear component distinct, so a misordered Voigt slot fails. The same increments,
// small enough to stay elastic, through the Johnson-Cook routine give the same stress.
void elasticMaterial() {
  const double E = 210000.0, nu = 0.30, dt = 2.5e-5;
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  VumatPdCorrespondenceConfig c;
  c.material.cmname = "ELASTIC";
  c.material.pluginPath = "elastic_vumat_plugin.so";
  c.material.nstatev = 1;
  c.material.props = {E, nu};
  c.material.density = 7800;
  c.initialState = {4.5};
  VumatPdCorrespondenceConfig j = c;
  j.material.cmname = "JOHNSONCOOK";
  j.material.pluginPath = "jc_vumat_plugin.so";
  j.material.nstatev = 6;
  j.material.props = {E, nu, 350.0, 500.0, 0.025, 0.36, 1.0, 1.0, 293.0, 1793.0, 7800.0,
                      0.9, 500.0, 1.0};
  j.initialState = {};
  VumatPdCorrespondenceMaterial el(c), jcm(j);
  const std::vector<Tensor3> path = {
      eye(),
      {1 + 1.0e-4, 2.0e-5, -1.5e-5, 3.5e-5, 1 - 3.0e-5, 0.5e-5, -2.5e-5, 4.0e-5, 1 + 2.0e-5},
      {1 + 1.8e-4, 3.0e-5, -1.0e-5, 6.5e-5, 1 - 5.0e-5, 1.5e-5, -4.0e-5, 7.0e-5, 1 + 3.0e-5}};
  MaterialState se, sj;
  el.initialize({path[0]}, se);
  jcm.initialize({path[0]}, sj);
  jcm.update({path[0]}, sj, dt, 0, 0); // the JC routine's first call is its init branch
  Tensor3 sbar{}, Vn = eye(), Rn = eye();
  for (int step = 1; step <= 2; ++step) {
    const auto kin = updateCorotationalKinematics(path[step - 1], path[step], Vn, Rn, dt);
    double tr = 0;
    for (int a = 0; a < 3; ++a)
      tr += kin.unrotatedRateOfDeformation[4 * a] * dt;
    for (int q = 0; q < 9; ++q)
      sbar[q] += 2 * mu * kin.unrotatedRateOfDeformation[q] * dt + (q % 4 == 0 ? lam * tr : 0);
    const Tensor3 expect = rotateCorotationalStressToSpatial(sbar, kin.rotationNP1);
    Vn = kin.leftStretchNP1;
    Rn = kin.rotationNP1;
    const auto out = el.update({path[step]}, se, dt, step * dt, step * dt);
    double scale = 0;
    for (double v : expect)
      scale = std::max(scale, std::abs(v));
    require(maxDiff(out.rawStress[0], expect) <= 1e-10 * scale, "elastic stress, Hooke's law");
    require(maxDiff(se.rawStressN[0], sbar) <= 1e-10 * scale, "elastic corotational history");
    require(se.vumatStateN[0].size() == 1 && se.vumatStateN[0][0] == 4.5, "state carried");
    require(se.temperatureN[0] == 293.0 && out.internalEnergy[0] == 0.0 &&
                out.inelasticEnergy[0] == 0.0,
            "temperature and energies carried");
    const auto oj = jcm.update({path[step]}, sj, dt, step * dt, step * dt);
    require(sj.vumatStateN[0][1] == 0.0, "the JC increments stay elastic");
    require(maxDiff(oj.rawStress[0], out.rawStress[0]) <= 1e-12 * scale,
            "elastic routine equals the JC routine below yield");
  }
}
} // namespace

int main() {
  const std::pair<const char *, std::function<void()>> tests[] = {
      {"neighbor lattice", neighborLattice}, {"correspondence sampling", correspondenceSampling},
      {"corotational", corotational},         {"outer damage", outerDamage},
      {"integrator", integrator},             {"material", material},
      {"elastic material", elasticMaterial}};
  for (const auto &t : tests) {
    std::printf("[class-api] %s\n", t.first);
    std::fflush(stdout);
    try {
      t.second();
    } catch (const std::exception &e) {
      std::printf("FAIL %s threw: %s\n", t.first, e.what());
      ++failures;
    }
  }
  if (failures)
    std::printf("class API tests: %d failure(s)\n", failures);
  else
    std::printf("class API tests passed\n");
  return failures ? 1 : 0;
}
