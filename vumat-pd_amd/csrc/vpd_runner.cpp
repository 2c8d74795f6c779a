// The JSON case runner on the GPU: reads the reference's case and material files
// (src/config/DemoConfig.cpp, src/config/MaterialConfigJson.cpp) and writes the reference's
// outputs.
// - Boundary-value cases (src/host/BoundaryValueCaseRunner.cpp): every material runs through
//   vpd_step with the fused run diagnostics; per step only one 16-double diagnostics record
//   crosses to the host (in batches).
// - Affine demo cases (src/host/DemoCaseRunner.cpp): the prescribed affine displacement, F,
//   the material update (Johnson-Cook or the linear elastic VUMAT), damage and the force
//   density run on the device each step; the small per-step summaries are formed on the host
//   in the reference's expressions.
#include "vpd.h"
#include "vpd_json.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using vpdjson::Value;

// ---- configuration (the reference's DemoCaseConfig and VumatMaterialConfig fields) --------

struct MaterialFile { // VumatMaterialConfig (include/VumatTypes.hpp), MaterialConfigJson.cpp:62-110
  std::string name, cmname, symbolName, stressFrame, pluginPath;
  int nstatev = 0;
  std::vector<double> props;
  double density = 0.0;
};

struct CaseMaterial { // DemoMaterialConfig (include/DemoConfig.hpp:83-96)
  std::string label;
  MaterialFile material;
  double elasticE = 0.0, elasticNu = 0.0, referenceTemperature = 293.0;
  std::vector<double> initialState;
  int eqpsIndex = -1;
  std::vector<int> indicatorIndices;
  bool indicatorMin = false, indicatorOneMinus = false;
  bool reference = false;
};

struct CaseConfig { // DemoCaseConfig (include/DemoConfig.hpp:98-113)
  std::string name, type, scope, classification;
  std::string unitLength = "mm", unitTime = "ms", unitForce = "N", unitStress = "MPa",
              unitMassDensity = "N*ms^2/mm^4";
  std::string geometryType = "patch";
  int nx = 1, ny = 1, nz = 1;
  double spacing = 1.0, horizon = 1.0;
  int steps = 1;
  double dt = 1.0e-6, totalTime = 0.0;
  std::string bcType = "affine_displacement", supportRegion = "left", loadRegion = "right",
              component = "x", bcMode = "uniaxial_xx";
  double maxStrain = 0.0;
  double amplitude = 0.0, riseTime = 0.0;
  int regionLayers = 1;
  bool planeStrain = false, rigidBodySuppression = false;
  double density = 0.0, energyTol = 1.0e-3, netForceTol = 1.0e-8, minPlastic = 0.0,
         kineticLimit = 1.0;
  int outputEvery = 1, snapshotEvery = 0;
  std::string damageMode = "none";
  double damageInitiation = 0.0, damageFailure = 1.0;
  std::string outputDir;
  double axialMin = std::numeric_limits<double>::quiet_NaN();
  double axialMax = std::numeric_limits<double>::quiet_NaN();
  // ValidationToleranceConfig (include/DemoConfig.hpp:72-76) and the checks list.
  double tolDefgrad = 1.0e-12, tolRefStress = 1.0e-6, tolNetForce = 1.0e-7;
  std::vector<std::string> checks;
  std::vector<CaseMaterial> materials;
};

std::string join_path(const std::string &dir, const std::string &file) {
  if (dir.empty())
    return file;
  return (dir.back() == '/' || dir.back() == '\\') ? dir + file : dir + "/" + file;
}

// MaterialConfigJson.cpp:62-110.
MaterialFile load_material_file(const std::string &path, const std::string &pluginDir) {
  using namespace vpdjson;
  const Value root = load_file(path);
  if (get_string(root, "schemaVersion", false, "1.0") != "1.0")
    throw std::runtime_error("Unsupported VUMAT material schemaVersion: " +
                             get_string(root, "schemaVersion", false, "1.0"));
  MaterialFile m;
  m.name = get_string(root, "name");
  m.cmname = get_string(root, "cmname");
  const std::string convention = get_string(root, "tensorConvention", false, "legacy_vumat_pd");
  m.stressFrame = get_string(root, "stressFrame", false, "corotational");
  if (convention != "legacy_vumat_pd")
    throw std::runtime_error("Unsupported tensorConvention; explicit mapping is required.");
  if (m.stressFrame != "corotational" && m.stressFrame != "spatial")
    throw std::runtime_error("stressFrame must be corotational or spatial.");
  m.symbolName =
      get_string(root, "symbolName", false, get_string(root, "symbol", false, "vumat_adapter_call"));
  const int ndir = get_int(root, "ndir", false, 3), nshr = get_int(root, "nshr", false, 3);
  m.nstatev = get_int(root, "nstatev");
  const int nfieldv = get_int(root, "nfieldv", false, 1);
  m.props = get_double_array(root, "props");
  m.density = get_double(root, "density");
  if (ndir != 3 || nshr != 3)
    throw std::runtime_error("The current VUMAT adapter requires ndir=3 and nshr=3.");
  if (m.nstatev < 0 || nfieldv < 0 || m.density <= 0.0)
    throw std::runtime_error("Invalid VUMAT dimensions or density in: " + path);
  const std::string explicitPath = get_string(root, "pluginPath", false, "");
  std::string plugin = explicitPath;
  if (plugin.empty()) {
    plugin = get_string(root, "plugin", false, "");
    if (plugin.empty())
      throw std::runtime_error("Material config requires either pluginPath or plugin.");
    if (plugin.find('/') == std::string::npos && plugin.find('\\') == std::string::npos &&
        plugin.find('.') == std::string::npos)
      plugin += ".so";
  }
  m.pluginPath = (!plugin.empty() && plugin[0] == '/') ? plugin : join_path(pluginDir, plugin);
  return m;
}

std::vector<int> int_array(const Value &o, const std::string &name) {
  const Value *v = vpdjson::find(o, name);
  if (!v)
    return {};
  if (v->kind != Value::Array)
    throw std::runtime_error("Expected integer array: " + name);
  std::vector<int> out;
  for (const Value &e : v->array) {
    if (e.kind != Value::Number || !std::isfinite(e.number) || std::trunc(e.number) != e.number ||
        e.number < 0 || e.number > std::numeric_limits<int>::max())
      throw std::runtime_error("Expected nonnegative integer entries in " + name);
    out.push_back(static_cast<int>(e.number));
  }
  return out;
}

// DemoConfig.cpp:206-281.
CaseMaterial load_case_material(const Value &e, const std::string &pluginDir) {
  using namespace vpdjson;
  CaseMaterial m;
  m.label = get_string(e, "name");
  const std::string path = get_string(
      e, "materialConfigPath", false,
      get_string(e, "configPath", false, get_string(e, "material_config_path", false, "")));
  if (path.empty())
    throw std::runtime_error("Material entry requires materialConfigPath or configPath.");
  m.material = load_material_file(path, pluginDir);
  if (m.material.stressFrame != "corotational")
    throw std::runtime_error(
        "JSON case hosts require corotational stress; use the dedicated spatial demo.");
  m.reference = get_bool(e, "reference", false, false);
  m.elasticE = get_double(e, "elasticE", false, 0.0);
  m.elasticNu = get_double(e, "elasticNu", false, 0.0);
  m.referenceTemperature = get_double(e, "referenceTemperature", false, m.referenceTemperature);
  if (get_string(e, "kinematicsMode", false, "corotational") != "corotational")
    throw std::runtime_error("Unsupported kinematicsMode for material: " + m.label);
  if (const Value *init = find(e, "initialState")) {
    if (init->kind != Value::Array)
      throw std::runtime_error("Expected numeric array: initialState");
    for (const Value &x : init->array) {
      if (x.kind != Value::Number)
        throw std::runtime_error("Expected numeric entries in initialState");
      m.initialState.push_back(x.number);
    }
  }
  const int nstatev = m.material.nstatev;
  if (!m.initialState.empty() && m.initialState.size() != static_cast<size_t>(nstatev))
    throw std::runtime_error("initialState length does not match nstatev.");
  m.eqpsIndex = get_int(e, "equivalentPlasticStrainIndex", false, -1);
  if (m.eqpsIndex < -1 || m.eqpsIndex >= nstatev)
    throw std::runtime_error("equivalentPlasticStrainIndex is outside nstatev for material: " +
                             m.label);
  const Value *over = find(e, "initialStateOverrides");
  if (over && over->kind == Value::Object) {
    if (m.initialState.empty())
      m.initialState.assign(static_cast<size_t>(nstatev), 0.0);
    for (const auto &item : over->object) {
      const std::string &key = item.first;
      const bool digits = !key.empty() && std::all_of(key.begin(), key.end(), [](char c) {
        return c >= '0' && c <= '9';
      });
      if (!digits)
        throw std::runtime_error("Invalid initialStateOverrides index for material: " + m.label);
      const long index = std::stol(key);
      if (index < 0 || index >= nstatev || item.second.kind != Value::Number)
        throw std::runtime_error("Invalid initialStateOverrides entry for material: " + m.label);
      m.initialState[static_cast<size_t>(index)] = item.second.number;
    }
  }
  const Value *ind = find(e, "internalIndicator");
  if (ind && ind->kind == Value::Object) {
    m.indicatorIndices = int_array(*ind, "indices");
    const std::string red = get_string(*ind, "reduction", false, "max");
    const std::string tr = get_string(*ind, "transform", false, "identity");
    if (red != "min" && red != "max")
      throw std::runtime_error("internalIndicator.reduction must be min or max.");
    if (tr != "identity" && tr != "one_minus")
      throw std::runtime_error("internalIndicator.transform must be identity or one_minus.");
    for (int k : m.indicatorIndices)
      if (k < 0 || k >= nstatev)
        throw std::runtime_error("internalIndicator index is outside nstatev for material: " +
                                 m.label);
    m.indicatorMin = red == "min";
    m.indicatorOneMinus = tr == "one_minus";
  }
  return m;
}

// DemoConfig.cpp:17-197 and the checks of :285-380 that apply to boundary cases.
CaseConfig load_case(const std::string &path, const std::string &pluginDir) {
  using namespace vpdjson;
  const Value root = load_file(path);
  const std::string schema = get_string(root, "schemaVersion", false, "1.0");
  if (schema != "1.0")
    throw std::runtime_error("Unsupported VUMAT-PD case schemaVersion: " + schema);
  CaseConfig c;
  c.name = get_string(root, "name");
  c.type = get_string(root, "type", false, "vumat_pd_demo");
  c.scope = get_string(root, "scope", false, "");
  c.classification = get_string(root, "classification", false, "verification");
  if (const Value *u = find(root, "units"); u && u->kind == Value::Object) {
    c.unitLength = get_string(*u, "length", false, c.unitLength);
    c.unitTime = get_string(*u, "time", false, c.unitTime);
    c.unitForce = get_string(*u, "force", false, c.unitForce);
    c.unitStress = get_string(*u, "stress", false, c.unitStress);
    c.unitMassDensity = get_string(*u, "massDensity", false, c.unitMassDensity);
  }
  const Value *geo = find(root, "geometry");
  if (!geo || geo->kind != Value::Object)
    geo = find(root, "grid");
  if (geo && geo->kind == Value::Object) {
    c.geometryType = get_string(*geo, "type", false, c.geometryType);
    c.nx = get_int(*geo, "nx", false, c.nx);
    c.ny = get_int(*geo, "ny", false, c.ny);
    c.nz = get_int(*geo, "nz", false, c.nz);
    c.spacing = get_double(*geo, "spacing", false, c.spacing);
    c.horizon = get_double(*geo, "horizon", false, c.horizon);
  }
  const Value *ti = find(root, "time_integration");
  const Value &tsrc = (ti && ti->kind == Value::Object) ? *ti : root;
  c.steps = get_int(tsrc, "steps", false, c.steps);
  c.dt = get_double(tsrc, "dt", false, c.dt);
  c.totalTime = get_double(tsrc, "totalTime", false, c.totalTime);
  if (const Value *bc = find(root, "boundary_conditions"); bc && bc->kind == Value::Object) {
    c.bcType = get_string(*bc, "type", false, c.bcType);
    c.bcMode = get_string(*bc, "mode", false, c.bcMode);
    c.maxStrain = get_double(*bc, "maxStrain", false, c.maxStrain);
    c.amplitude = get_double(*bc, "amplitude", false, c.amplitude);
    c.riseTime = get_double(*bc, "riseTime", false, c.riseTime);
    (void)get_bool(*bc, "transverseConstraint", false, false);
    c.supportRegion = get_string(*bc, "supportRegion", false, c.supportRegion);
    c.loadRegion = get_string(*bc, "loadRegion", false, c.loadRegion);
    c.component = get_string(*bc, "component", false, c.component);
    c.regionLayers = get_int(*bc, "regionLayers", false, c.regionLayers);
    c.planeStrain = get_bool(*bc, "planeStrain", false, c.planeStrain);
    c.rigidBodySuppression = get_bool(*bc, "rigidBodySuppression", false, false);
  }
  if (const Value *d = find(root, "dynamics"); d && d->kind == Value::Object) {
    c.density = get_double(*d, "density", false, c.density);
    c.energyTol = get_double(*d, "energyToleranceRelative", false, c.energyTol);
    c.netForceTol = get_double(*d, "netForceTolerance", false, c.netForceTol);
    c.outputEvery = get_int(*d, "outputEvery", false, c.outputEvery);
    c.snapshotEvery = get_int(*d, "snapshotEvery", false, c.snapshotEvery);
    c.minPlastic = get_double(*d, "minimumPlasticStrain", false, c.minPlastic);
    c.kineticLimit = get_double(*d, "kineticToInternalLimit", false, c.kineticLimit);
  }
  if (const Value *d = find(root, "damage"); d && d->kind == Value::Object) {
    const std::string model = get_string(*d, "model", false, "");
    if (model == "critical_stretch" || model == "Critical Stretch")
      c.damageMode = "critical_stretch";
    else if (!model.empty())
      c.damageMode = model;
    c.damageInitiation = get_double(*d, "initiation", false, c.damageInitiation);
    if (c.damageInitiation <= 0.0)
      c.damageInitiation = get_double(*d, "criticalStretch", false, c.damageInitiation);
    c.damageFailure = get_double(*d, "failure", false, c.damageFailure);
    (void)get_bool(*d, "deletionEnabled", false, false);
  }
  c.outputDir = "validation_cases/output/" + c.name;
  if (const Value *o = find(root, "output"); o && o->kind == Value::Object)
    c.outputDir = get_string(*o, "directory", false, c.outputDir);
  if (const Value *ob = find(root, "observation")) {
    if (ob->kind != Value::Object)
      throw std::runtime_error("observation must be a JSON object.");
    c.axialMin = get_double(*ob, "axialMin", false, c.axialMin);
    c.axialMax = get_double(*ob, "axialMax", false, c.axialMax);
  }
  if (const Value *t = find(root, "tolerances"); t && t->kind == Value::Object) {
    c.tolDefgrad = get_double(*t, "deformationGradientAbsolute", false, c.tolDefgrad);
    c.tolRefStress = get_double(*t, "referenceStressAbsolute", false, c.tolRefStress);
    c.tolNetForce = get_double(*t, "netForceAbsolute", false, c.tolNetForce);
  }
  if (const Value *ch = find(root, "checks"); ch && ch->kind == Value::Array)
    for (const Value &item : ch->array)
      if (item.kind == Value::String)
        c.checks.push_back(item.string);
  const Value *mats = find(root, "materials");
  if (!mats || mats->kind != Value::Array)
    throw std::runtime_error("Demo case requires a materials array.");
  for (const Value &e : mats->array) {
    if (e.kind != Value::Object)
      throw std::runtime_error("Each materials entry must be a JSON object.");
    c.materials.push_back(load_case_material(e, pluginDir));
  }
  if (c.totalTime <= 0.0)
    c.totalTime = c.dt * static_cast<double>(c.steps);
  if (c.type != "vumat_pd_demo_case" && c.type != "vumat_pd_boundary_case" &&
      c.type != "vumat_material_point_case")
    throw std::runtime_error("Unsupported case type: " + c.type);
  if (c.spacing <= 0.0 || c.horizon <= 0.0)
    throw std::runtime_error("PD geometry requires positive spacing and horizon.");
  if (c.geometryType != "patch")
    throw std::runtime_error("Unsupported geometry type: " + c.geometryType);
  if (c.nx < 2 || c.ny < 2 || c.nz < 2)
    throw std::runtime_error("PD patch geometry requires nx/ny/nz >= 2.");
  if (c.steps < 1 || !std::isfinite(c.dt) || c.dt <= 0.0)
    throw std::runtime_error("time_integration requires steps >= 1 and dt > 0.");
  const double expected = c.dt * c.steps;
  if (!std::isfinite(expected) || !std::isfinite(c.totalTime))
    throw std::runtime_error("time_integration duration must be finite.");
  if (std::abs(c.totalTime - expected) > 1.0e-12 * std::max(1.0, expected))
    throw std::runtime_error("time_integration.totalTime must equal steps * dt.");
  if (c.type != "vumat_pd_boundary_case") { // demo and material-point cases
    if (c.bcType != "affine_displacement" ||
        (c.bcMode != "uniaxial_xx" && c.bcMode != "simple_shear_xy" &&
         c.bcMode != "plane_strain_x_uniaxial_stress"))
      throw std::runtime_error("Unsupported boundary_conditions type or mode.");
    if (c.materials.empty())
      throw std::runtime_error("Demo case requires at least one material.");
    return c;
  }
  if (c.bcType != "regional_smooth_displacement" || c.amplitude == 0.0 || c.riseTime <= 0.0 ||
      c.regionLayers < 1 || !c.rigidBodySuppression)
    throw std::runtime_error(
        "Boundary-value case requires regional smooth loading and rigid-body suppression.");
  const bool xCase = c.component == "x" && c.supportRegion == "left" && c.loadRegion == "right";
  const bool yCase = c.component == "y" && c.supportRegion == "bottom" && c.loadRegion == "top";
  if (!xCase && !yCase)
    throw std::runtime_error(
        "Regional boundary case supports left/right-x or bottom/top-y loading.");
  if (c.nz < 3 || c.horizon / c.spacing < 1.5)
    throw std::runtime_error("PD boundary cases require nz>=3 and horizon/spacing>=1.5.");
  if (c.density <= 0.0 || c.outputEvery < 1 || c.snapshotEvery < 1 || c.energyTol <= 0.0 ||
      c.netForceTol <= 0.0 || c.kineticLimit <= 0.0)
    throw std::runtime_error("Invalid boundary-value dynamics configuration.");
  if (c.unitLength != "mm" || c.unitTime != "ms" || c.unitForce != "N" || c.unitStress != "MPa" ||
      c.unitMassDensity != "N*ms^2/mm^4")
    throw std::runtime_error(
        "Boundary-value cases currently require the documented mm-ms-N-MPa unit system.");
  if (c.materials.empty())
    throw std::runtime_error("Demo case requires at least one material.");
  return c;
}

// ---- output formatting (iostream's default float format at the stream's precision) -------

std::string num(double x, int precision) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*g", precision, x);
  return buf;
}

std::string json_number(double x) { return std::isfinite(x) ? num(x, 15) : "null"; }

struct RunSummary { // BoundaryValueCaseRunner.cpp:32-52
  std::string material;
  bool pass = false, finite = true;
  double runtimeSeconds = 0, maxNetInternalForce = 0, finalEnergyError = 0, maxEnergyError = 0,
         maxKineticToInternal = 0, finalKineticToInternal = 0, maxEqps = 0,
         maxInternalIndicator = 0, earlyElasticSlope = 0, analyticalPlaneStrainSlope = 0,
         earlyElasticRelativeError = 0, finalReaction = 0, finalNominalStress = 0,
         finalEngineeringStrain = 0, observationMin = 0, observationMax = 0;
};

void check(vpd_context *ctx, int status) {
  if (status != VPD_OK)
    throw std::runtime_error(vpd_last_error(ctx));
}

struct ContextDeleter {
  void operator()(vpd_context *c) const { vpd_destroy(c); }
};

double internal_indicator(const CaseMaterial &m, const double *state) {
  if (m.indicatorIndices.empty())
    return 0.0;
  double value = m.indicatorMin ? std::numeric_limits<double>::infinity()
                                : -std::numeric_limits<double>::infinity();
  for (int k : m.indicatorIndices)
    value = m.indicatorMin ? std::min(value, state[k]) : std::max(value, state[k]);
  return m.indicatorOneMinus ? 1.0 - value : value;
}

std::vector<char> region(const std::vector<double> &x, int64_t n, const char *side,
                         double width) {
  std::vector<uint8_t> mask(n);
  const int64_t count = vpd_select_boundary_region(x.data(), n, side, width, mask.data());
  if (count == -VPD_ERR_ARGUMENT && width > 0.0)
    throw std::runtime_error("Boundary-region selection returned no particles.");
  if (count < 0)
    throw std::runtime_error("Boundary-region width and particle set must be nonzero.");
  return std::vector<char>(mask.begin(), mask.end());
}

// meanCoordinate (BoundaryValueCaseRunner.cpp:85-93) over a mask, in index order.
double mean_of(const std::vector<double> &field, const std::vector<char> &mask, int comp) {
  double sum = 0.0;
  int64_t count = 0;
  for (size_t i = 0; i < mask.size(); ++i)
    if (mask[i]) {
      sum += field[3 * i + comp];
      ++count;
    }
  return sum / static_cast<double>(count);
}

// One boundary-value run: one material of one case (runMaterial, BoundaryValueCaseRunner.cpp:
// 139-307), with the particle work on the GPU. setup() is runMaterial's preamble (patch, bands,
// observation window and their checks); consume() is the host half of the per-step
// diagnostics (:240-305) for one step record; snapshot() writes writeSnapshot's rows
// (:114-136). The run's output rows are kept until the case's files are written.
struct BoundaryRun {
  const CaseConfig *cfg = nullptr;
  const CaseMaterial *mat = nullptr;
  int64_t n = 0;
  std::vector<double> x, vol, weight;
  std::vector<uint8_t> codes, latSets;
  int axis = 0, lateral = 1;
  double loadingSign = 1.0, gaugeLength = 0.0, lateralGauge = 0.0, interiorVolume = 0.0;
  double crossSectionArea = 0.0;
  int64_t countNeg = 0, countPos = 0;
  RunSummary summary;
  double internalEnergy = 0.0, externalWork = 0.0, previousPower = 0.0;
  double slopeX2 = 0.0, slopeXY = 0.0;
  std::string ts, snap;

  BoundaryRun(const CaseConfig &c, const CaseMaterial &m) : cfg(&c), mat(&m) {}

  void setup() {
    const CaseConfig &cf = *cfg;
    n = static_cast<int64_t>(cf.nx) * cf.ny * cf.nz;
    x.assign(3 * n, 0.0);
    vol.assign(n, 0.0);
    if (vpd_make_patch(cf.nx, cf.ny, cf.nz, cf.spacing, x.data(), vol.data()) != VPD_OK)
      throw std::runtime_error("Particle patch requires positive dimensions and spacing.");
    axis = cf.component == "y" ? 1 : 0;
    lateral = axis == 0 ? 1 : 0;
    const double width =
        static_cast<double>(cf.regionLayers - 1) * cf.spacing + 1.0e-9 * cf.spacing;
    const auto support = region(x, n, cf.supportRegion.c_str(), width);
    const auto loaded = region(x, n, cf.loadRegion.c_str(), width);
    codes.assign(n, 0);
    int64_t anchor = 0, pin = 0;
    if (vpd_regional_constraint_codes(x.data(), n, axis, cf.regionLayers, cf.spacing,
                                      cf.planeStrain ? 1 : 0, codes.data(), &anchor,
                                      &pin) != VPD_OK)
      throw std::runtime_error("Boundary-region selection returned no particles.");
    loadingSign = cf.amplitude > 0.0 ? 1.0 : -1.0;
    gaugeLength = std::abs(mean_of(x, loaded, axis) - mean_of(x, support, axis));
    const auto latNeg = region(x, n, lateral == 0 ? "left" : "bottom", 1.0e-9 * cf.spacing);
    const auto latPos = region(x, n, lateral == 0 ? "right" : "top", 1.0e-9 * cf.spacing);
    lateralGauge = std::abs(mean_of(x, latPos, lateral) - mean_of(x, latNeg, lateral));
    latSets.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) {
      countNeg += latNeg[i] ? 1 : 0;
      countPos += latPos[i] ? 1 : 0;
      latSets[i] = (latNeg[i] ? 1 : 0) | (latPos[i] ? 2 : 0);
    }
    // observationBounds / observationWeights (include/ObservationWindow.hpp:13-58).
    double lo = x[axis], hi = lo;
    for (int64_t i = 0; i < n; ++i) {
      lo = std::min(lo, x[3 * i + axis]);
      hi = std::max(hi, x[3 * i + axis]);
    }
    const double band = (cf.regionLayers - 1) * cf.spacing;
    const double obsA = std::isnan(cf.axialMin) ? lo + band : cf.axialMin;
    const double obsB = std::isnan(cf.axialMax) ? hi - band : cf.axialMax;
    if (!std::isfinite(obsA) || !std::isfinite(obsB) || obsA < lo || obsB > hi || obsA >= obsB)
      throw std::runtime_error(
          "Observation window must lie within the body and have positive length.");
    const double tol = 1e-8 * cf.spacing;
    weight.assign(n, 0.0);
    bool foundA = false, foundB = false;
    for (int64_t i = 0; i < n; ++i) {
      const double xi = x[3 * i + axis];
      const bool atA = std::abs(xi - obsA) < tol, atB = std::abs(xi - obsB) < tol;
      foundA = foundA || atA;
      foundB = foundB || atB;
      if (xi < obsA - tol || xi > obsB + tol)
        continue;
      const bool interiorEnd = (atA && obsA > lo + tol) || (atB && obsB < hi - tol);
      weight[i] = vol[i] * (interiorEnd ? 0.5 : 1.0);
    }
    if (!foundA || !foundB)
      throw std::runtime_error("Observation bounds must coincide with reference nodal planes.");
    for (int64_t i = 0; i < n; ++i)
      if (weight[i] > 0.0)
        interiorVolume += weight[i];
    const int crossCount = axis == 0 ? cf.ny : cf.nx;
    crossSectionArea =
        static_cast<double>((crossCount - 1) * (cf.nz - 1)) * cf.spacing * cf.spacing;
    // The material must be the device's Johnson-Cook VUMAT.
    const std::string plugin =
        std::filesystem::path(mat->material.pluginPath).filename().string();
    std::string cm = mat->material.cmname;
    std::transform(cm.begin(), cm.end(), cm.begin(),
                   [](unsigned char ch) { return std::toupper(ch); });
    if (cm.rfind("JOHNSONCOOK", 0) != 0 && plugin.rfind("jc_vumat_plugin", 0) != 0)
      throw std::runtime_error("Only the Johnson-Cook VUMAT runs on the device; material " +
                               mat->label + " uses " + plugin);
    summary.material = mat->label;
    summary.observationMin = obsA;
    summary.observationMax = obsB;
    summary.analyticalPlaneStrainSlope = mat->elasticE / (1.0 - mat->elasticNu * mat->elasticNu);
  }

  vpd_config context_config() const {
    const CaseConfig &cf = *cfg;
    vpd_config vc{};
    vc.horizon = cf.horizon;
    vc.density = cf.density;
    vc.nprops = static_cast<int>(mat->material.props.size());
    vc.props = mat->material.props.data();
    vc.nstatev = mat->material.nstatev;
    vc.initial_state = mat->initialState.empty() ? nullptr : mat->initialState.data();
    vc.reference_temperature = mat->referenceTemperature;
    vc.char_length = cf.spacing;
    vc.material_density = mat->material.density;
    if (cf.damageMode == "none") {
      vc.damage_mode = VPD_DAMAGE_NONE;
    } else if (cf.damageMode == "critical_stretch") {
      if (cf.damageInitiation <= 0.0)
        throw std::runtime_error("critical_stretch requires initiation > 0.");
      vc.damage_mode = VPD_DAMAGE_CRITICAL_STRETCH;
    } else {
      throw std::runtime_error("Unsupported damage model on the device: " + cf.damageMode);
    }
    vc.damage_initiation = cf.damageInitiation;
    vc.damage_failure = cf.damageFailure;
    vc.return_map = VPD_RETURN_MAP_BISECTION;
    return vc;
  }

  // The diagnostics configuration (weights and face sets of this run's particles).
  vpd_diag_config diag_config(int capacity) const {
    vpd_diag_config dc{};
    dc.load_axis = axis;
    dc.lateral_axis = lateral;
    dc.observation_weight = weight.data();
    dc.lateral_sets = latSets.data();
    dc.eqps_index = mat->eqpsIndex;
    dc.indicator_count = static_cast<int>(mat->indicatorIndices.size());
    if (dc.indicator_count > 8)
      throw std::runtime_error("At most 8 internal-indicator state indices run on the device.");
    for (int k = 0; k < dc.indicator_count; ++k)
      dc.indicator_indices[k] = mat->indicatorIndices[k];
    dc.indicator_min = mat->indicatorMin ? 1 : 0;
    dc.indicator_one_minus = mat->indicatorOneMinus ? 1 : 0;
    dc.capacity = capacity;
    return dc;
  }

  void ramp(int step, double out[3]) const {
    vpd_smooth_ramp(cfg->amplitude, cfg->riseTime, static_cast<double>(step) * cfg->dt, out);
  }

  bool snapshot_step(int step) const {
    return step % cfg->snapshotEvery == 0 || step == cfg->steps;
  }

  void consume(const double *r, int step) {
    const CaseConfig &cf = *cfg;
    const double dt = cf.dt;
    const double time = static_cast<double>(step) * dt;
    double rp[3];
    ramp(step, rp);
    internalEnergy += r[VPD_DIAG_INTERNAL_ENERGY];
    const double net = std::sqrt(r[VPD_DIAG_NET_FORCE_X] * r[VPD_DIAG_NET_FORCE_X] +
                                 r[VPD_DIAG_NET_FORCE_Y] * r[VPD_DIAG_NET_FORCE_Y] +
                                 r[VPD_DIAG_NET_FORCE_Z] * r[VPD_DIAG_NET_FORCE_Z]);
    const double reaction = r[VPD_DIAG_REACTION];
    const double averageStress = loadingSign * r[VPD_DIAG_OBSERVED_STRESS] / interiorVolume;
    const double averageAxialStrain = loadingSign * r[VPD_DIAG_OBSERVED_STRAIN] / interiorVolume;
    const double maxEqps = r[VPD_DIAG_MAX_EQPS];
    const double maxIndicator = r[VPD_DIAG_MAX_INDICATOR];
    summary.finite = summary.finite && r[VPD_DIAG_NONFINITE_MATERIAL] == 0.0 &&
                     r[VPD_DIAG_NONFINITE_FORCE] == 0.0;
    const double power = reaction * rp[1];
    externalWork += 0.5 * dt * (previousPower + power);
    previousPower = power;
    const double kinetic = r[VPD_DIAG_KINETIC];
    const double scale =
        std::max({std::abs(externalWork), std::abs(internalEnergy) + kinetic, 1.0e-12});
    const double energyError = std::abs(kinetic + internalEnergy - externalWork) / scale;
    const double kineticRatio = kinetic / std::max(std::abs(internalEnergy), 1.0e-12);
    const double engineeringStrain = loadingSign * rp[0] / gaugeLength;
    const double lateralStrain = (r[VPD_DIAG_LATERAL_POS] / static_cast<double>(countPos) -
                                  r[VPD_DIAG_LATERAL_NEG] / static_cast<double>(countNeg)) /
                                 lateralGauge;
    const double nominalStress = loadingSign * reaction / crossSectionArea;
    summary.maxNetInternalForce = std::max(summary.maxNetInternalForce, net);
    summary.maxEnergyError = std::max(summary.maxEnergyError, energyError);
    summary.maxKineticToInternal = std::max(summary.maxKineticToInternal, kineticRatio);
    summary.finalKineticToInternal = kineticRatio;
    summary.maxEqps = std::max(summary.maxEqps, maxEqps);
    summary.maxInternalIndicator = std::max(summary.maxInternalIndicator, maxIndicator);
    summary.finalEnergyError = energyError;
    summary.finalReaction = reaction;
    summary.finalNominalStress = nominalStress;
    summary.finalEngineeringStrain = engineeringStrain;
    summary.finite = summary.finite && std::isfinite(kinetic) && std::isfinite(internalEnergy) &&
                     std::isfinite(externalWork) && std::isfinite(energyError);
    if (engineeringStrain > 1.0e-6 && engineeringStrain <= 1.0e-3) {
      slopeX2 += engineeringStrain * engineeringStrain;
      slopeXY += engineeringStrain * averageStress;
    }
    if (step % cf.outputEvery == 0 || step == 1 || step == cf.steps) {
      const double vals[] = {time, rp[0], rp[1], reaction, engineeringStrain, lateralStrain,
                             nominalStress, averageAxialStrain, averageStress, maxEqps,
                             maxIndicator, kinetic, internalEnergy, externalWork, energyError,
                             net};
      ts += mat->label + ',' + std::to_string(step);
      for (double v : vals)
        ts += ',' + num(v, 16);
      ts += summary.finite ? ",1\n" : ",0\n";
    }
  }

  // The run's particles start at `offset` in the downloaded block fields.
  void snapshot(int step, int64_t offset, int nstatev, const double *u, const double *v,
                const double *stress, const double *strain, const double *statev,
                const double *dmg) {
    const double time = static_cast<double>(step) * cfg->dt;
    std::string line;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t j = offset + i;
      const double *x0 = &x[3 * i], *ui = &u[3 * j], *vi = &v[3 * j];
      const double *st = &statev[static_cast<size_t>(j) * nstatev];
      const double eqps = mat->eqpsIndex >= 0 ? st[mat->eqpsIndex] : 0.0;
      const double vals[] = {x0[0], x0[1], x0[2], ui[0], ui[1], ui[2],
                             x0[0] + ui[0], x0[1] + ui[1], x0[2] + ui[2], vi[0], vi[1], vi[2],
                             strain[9 * j + 4 * axis], stress[9 * j + 4 * axis], eqps,
                             internal_indicator(*mat, st), dmg[j]};
      line = mat->label + ',' + std::to_string(step) + ',' + num(time, 16) + ',' +
             std::to_string(i);
      for (double val : vals)
        line += ',' + num(val, 16);
      line += '\n';
      snap += line;
    }
  }

  void finish(double runtimeSeconds) {
    const CaseConfig &cf = *cfg;
    summary.earlyElasticSlope = slopeX2 > 0.0 ? slopeXY / slopeX2 : 0.0;
    summary.earlyElasticRelativeError =
        summary.analyticalPlaneStrainSlope > 0.0
            ? std::abs(summary.earlyElasticSlope - summary.analyticalPlaneStrainSlope) /
                  summary.analyticalPlaneStrainSlope
            : 0.0;
    summary.runtimeSeconds = runtimeSeconds;
    summary.pass = summary.finite && summary.maxNetInternalForce <= cf.netForceTol &&
                   summary.finalEnergyError <= cf.energyTol &&
                   summary.finalKineticToInternal <= cf.kineticLimit &&
                   summary.maxEqps >= cf.minPlastic;
  }
};

// Fields of the whole block at a snapshot step.
struct BlockFields {
  std::vector<double> u, v, stress, strain, statev, dmg;
  void download(vpd_context *ctx, int64_t n, int nstatev) {
    u.resize(3 * n);
    v.resize(3 * n);
    stress.resize(9 * n);
    strain.resize(9 * n);
    dmg.resize(n);
    statev.resize(static_cast<size_t>(nstatev) * n);
    check(ctx, vpd_download(ctx, VPD_FIELD_DISPLACEMENT, u.data()));
    check(ctx, vpd_download(ctx, VPD_FIELD_VELOCITY, v.data()));
    check(ctx, vpd_download(ctx, VPD_FIELD_STRESS, stress.data()));
    check(ctx, vpd_download(ctx, VPD_FIELD_STRAIN, strain.data()));
    check(ctx, vpd_download(ctx, VPD_FIELD_STATEV, statev.data()));
    check(ctx, vpd_download(ctx, VPD_FIELD_DAMAGE_INDEX, dmg.data()));
  }
};

// runMaterial for one material alone: steps through vpd_step with the fused diagnostics.
void run_material(BoundaryRun &run) {
  const auto started = std::chrono::steady_clock::now();
  const CaseConfig &cfg = *run.cfg;
  run.setup();
  const int64_t n = run.n;
  const vpd_config vc = run.context_config();
  vpd_context *raw = nullptr;
  if (vpd_create(&vc, &raw) != VPD_OK)
    throw std::runtime_error(vpd_last_error(nullptr));
  std::unique_ptr<vpd_context, ContextDeleter> owner(raw);
  vpd_context *ctx = raw;
  check(ctx, vpd_set_particles(ctx, run.x.data(), run.vol.data(), n));
  int64_t nnz = 0;
  check(ctx, vpd_build_neighbors(ctx, &nnz));
  check(ctx, vpd_set_constraints(ctx, run.codes.data()));
  check(ctx, vpd_initialize_shape(ctx));
  check(ctx, vpd_initialize_material(ctx));
  check(ctx, vpd_initialize_dynamics(ctx));
  double ramp0[3];
  run.ramp(0, ramp0);
  check(ctx, vpd_apply_constraints(ctx, ramp0));
  const vpd_diag_config dc = run.diag_config(std::min(cfg.steps, 2048));
  check(ctx, vpd_enable_diagnostics(ctx, &dc));

  std::vector<double> records(static_cast<size_t>(dc.capacity) * VPD_DIAG_STRIDE);
  int processed = 0; // steps whose record has been consumed
  auto drain = [&] {
    int count = 0;
    check(ctx, vpd_fetch_diagnostics(ctx, records.data(), dc.capacity, &count));
    for (int k = 0; k < count; ++k)
      run.consume(&records[static_cast<size_t>(k) * VPD_DIAG_STRIDE], ++processed);
  };
  BlockFields fields;
  int pending = 0;
  for (int step = 1; step <= cfg.steps; ++step) {
    const bool snapshot = run.snapshot_step(step);
    vpd_step_args a{};
    a.dt = cfg.dt;
    a.time = static_cast<double>(step) * cfg.dt;
    a.step_time = (step - 1) * cfg.dt;
    run.ramp(step, a.ramp);
    // The next predictor is fused into this step unless the state after this step is read.
    a.fuse_next = snapshot ? 0 : 1;
    if (a.fuse_next)
      run.ramp(step + 1, a.ramp_next); // the time the next step recomputes, bit for bit
    a.want_fields = snapshot ? 1 : 0;
    if (pending == dc.capacity) {
      drain();
      pending = 0;
    }
    check(ctx, vpd_step(ctx, &a));
    ++pending;
    if (!snapshot)
      continue;
    drain();
    pending = 0;
    fields.download(ctx, n, run.mat->material.nstatev);
    run.snapshot(step, 0, run.mat->material.nstatev, fields.u.data(), fields.v.data(),
                 fields.stress.data(), fields.strain.data(), fields.statev.data(),
                 fields.dmg.data());
  }
  if (pending)
    drain();
  run.finish(std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count());
}

// The materials of several boundary cases as specimens of one particle block
// (vpd_set_specimens with half-cosine band tables, vpd_set_specimen_bands), stepped together
// with per-specimen diagnostics. Each run's rows equal those of run_material on the same case,
// bit for bit: every specimen keeps its own neighbour rows, time step, band values, step count
// and diagnostics chunks. The runs must share the device context's settings (material,
// solver density, damage) and the diagnostics configuration (load axis, state indices).
void run_batch(std::vector<BoundaryRun> &runs) {
  const auto started = std::chrono::steady_clock::now();
  for (auto &r : runs)
    r.setup();
  const BoundaryRun &head = runs.front();
  const vpd_config vc = head.context_config();
  for (const auto &r : runs) {
    const vpd_config o = r.context_config();
    const CaseMaterial &a = *head.mat, &b = *r.mat;
    if (o.density != vc.density || o.damage_mode != vc.damage_mode ||
        o.damage_initiation != vc.damage_initiation || o.damage_failure != vc.damage_failure ||
        a.material.props != b.material.props || a.material.nstatev != b.material.nstatev ||
        a.material.density != b.material.density || a.initialState != b.initialState ||
        a.referenceTemperature != b.referenceTemperature || a.eqpsIndex != b.eqpsIndex ||
        a.indicatorIndices != b.indicatorIndices || a.indicatorMin != b.indicatorMin ||
        a.indicatorOneMinus != b.indicatorOneMinus || r.axis != head.axis)
      throw std::runtime_error("Cases " + head.cfg->name + " and " + r.cfg->name +
                               " cannot share a block: material, solver density, damage, "
                               "load axis and state indices must agree.");
  }
  const int count = static_cast<int>(runs.size());
  std::vector<int64_t> first(count + 1, 0);
  for (int s = 0; s < count; ++s)
    first[s + 1] = first[s] + runs[s].n;
  const int64_t n = first[count];
  std::vector<double> x, vol, weight;
  std::vector<uint8_t> codes, latSets;
  std::vector<int32_t> rowptr(1, 0), col;
  for (auto &r : runs) { // each specimen's own neighbour rows, at its own horizon
    x.insert(x.end(), r.x.begin(), r.x.end());
    vol.insert(vol.end(), r.vol.begin(), r.vol.end());
    weight.insert(weight.end(), r.weight.begin(), r.weight.end());
    codes.insert(codes.end(), r.codes.begin(), r.codes.end());
    latSets.insert(latSets.end(), r.latSets.begin(), r.latSets.end());
    vpd_config rc = r.context_config();
    vpd_context *tmp = nullptr;
    if (vpd_create(&rc, &tmp) != VPD_OK)
      throw std::runtime_error(vpd_last_error(nullptr));
    std::unique_ptr<vpd_context, ContextDeleter> hold(tmp);
    check(tmp, vpd_set_particles(tmp, r.x.data(), r.vol.data(), r.n));
    int64_t nnz = 0;
    check(tmp, vpd_build_neighbors(tmp, &nnz));
    std::vector<int32_t> rp(r.n + 1), cl(std::max<int64_t>(1, nnz));
    check(tmp, vpd_get_neighbors(tmp, rp.data(), cl.data()));
    const int32_t base = static_cast<int32_t>(x.size() / 3 - r.n), offset = rowptr.back();
    for (int64_t i = 1; i <= r.n; ++i)
      rowptr.push_back(offset + rp[i]);
    for (int64_t k = 0; k < nnz; ++k)
      col.push_back(base + cl[k]);
  }
  vpd_context *raw = nullptr;
  if (vpd_create(&vc, &raw) != VPD_OK)
    throw std::runtime_error(vpd_last_error(nullptr));
  std::unique_ptr<vpd_context, ContextDeleter> owner(raw);
  vpd_context *ctx = raw;
  check(ctx, vpd_set_particles(ctx, x.data(), vol.data(), n));
  if (col.empty())
    col.push_back(0);
  check(ctx, vpd_set_neighbors(ctx, rowptr.data(), col.data(), rowptr.back()));
  check(ctx, vpd_set_constraints(ctx, codes.data()));
  check(ctx, vpd_initialize_shape(ctx));
  check(ctx, vpd_initialize_material(ctx));
  check(ctx, vpd_initialize_dynamics(ctx));
  std::vector<double> dts(count), vel(count, 0.0), table;
  std::vector<int32_t> last(count);
  std::vector<int64_t> offsets(count + 1, 0);
  int maxSteps = 0;
  for (int s = 0; s < count; ++s) {
    const BoundaryRun &r = runs[s];
    dts[s] = r.cfg->dt;
    last[s] = r.cfg->steps;
    maxSteps = std::max(maxSteps, r.cfg->steps);
    for (int k = 0; k <= r.cfg->steps; ++k) { // the ramp at step·dt, as run_material has it
      double rp[3];
      r.ramp(k, rp);
      table.insert(table.end(), rp, rp + 3);
    }
    offsets[s + 1] = offsets[s] + r.cfg->steps + 1;
  }
  check(ctx, vpd_set_specimens(ctx, count, first.data(), dts.data(), vel.data(), last.data()));
  check(ctx, vpd_set_specimen_bands(ctx, offsets.data(), table.data()));
  vpd_diag_config dc = head.diag_config(std::min(maxSteps, 1024));
  dc.observation_weight = weight.data();
  dc.lateral_sets = latSets.data();
  check(ctx, vpd_enable_diagnostics(ctx, &dc));

  std::vector<double> records(static_cast<size_t>(dc.capacity) * count * VPD_DIAG_STRIDE);
  int processed = 0;
  auto drain = [&] {
    int steps = 0;
    check(ctx, vpd_fetch_diagnostics(ctx, records.data(), dc.capacity, &steps));
    for (int k = 0; k < steps; ++k) {
      ++processed;
      for (int s = 0; s < count; ++s)
        if (processed <= runs[s].cfg->steps)
          runs[s].consume(&records[(static_cast<size_t>(k) * count + s) * VPD_DIAG_STRIDE],
                          processed);
    }
  };
  const int nstatev = head.mat->material.nstatev;
  BlockFields fields;
  int pending = 0;
  for (int step = 1; step <= maxSteps; ++step) {
    bool snapshot = false, midRun = false;
    for (const auto &r : runs)
      if (step <= r.cfg->steps && r.snapshot_step(step)) {
        snapshot = true;
        midRun = midRun || step < r.cfg->steps;
      }
    if (pending == dc.capacity) {
      drain();
      pending = 0;
    }
    // A specimen's last step is never followed by a prediction; a mid-run snapshot needs
    // the end-of-step state, so that step is left unfused.
    check(ctx, vpd_step_specimens(ctx, step, midRun ? 0 : 1, snapshot ? 1 : 0));
    ++pending;
    if (!snapshot)
      continue;
    drain();
    pending = 0;
    fields.download(ctx, n, nstatev);
    for (int s = 0; s < count; ++s)
      if (step <= runs[s].cfg->steps && runs[s].snapshot_step(step))
        runs[s].snapshot(step, first[s], nstatev, fields.u.data(), fields.v.data(),
                         fields.stress.data(), fields.strain.data(), fields.statev.data(),
                         fields.dmg.data());
  }
  if (pending)
    drain();
  const double runtime =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
  for (auto &r : runs)
    r.finish(runtime);
}

// runBoundaryValueCase's outputs (BoundaryValueCaseRunner.cpp:341-440) for a case whose
// material runs are done: timeseries.csv, field_snapshots.csv, metadata.json and the summary
// line per material. Returns 0 when every material passes, 2 otherwise.
int write_boundary_case(const CaseConfig &cfg, const std::vector<const BoundaryRun *> &runs) {
  std::filesystem::create_directories(cfg.outputDir);
  std::ofstream ts(cfg.outputDir + "/timeseries.csv");
  std::ofstream snap(cfg.outputDir + "/field_snapshots.csv");
  if (!ts || !snap)
    throw std::runtime_error("Failed to open boundary-value output files.");
  ts << "material,step,time_ms,load_displacement_mm,load_velocity_mm_per_ms,"
        "reaction_force_N,engineering_strain,lateral_strain,nominal_stress_MPa,"
        "average_axial_strain,average_axial_stress_MPa,max_eqps,max_internal_indicator,"
        "kinetic_energy_Nmm,internal_energy_Nmm,external_work_Nmm,relative_energy_error,"
        "net_internal_force_N,finite\n";
  snap << "material,step,time_ms,particle,x0_mm,y0_mm,z0_mm,ux_mm,uy_mm,uz_mm,"
          "x_mm,y_mm,z_mm,vx_mm_per_ms,vy_mm_per_ms,vz_mm_per_ms,axial_strain,"
          "axial_stress_MPa,eqps,internal_indicator,simple_damage\n";
  const long particles = static_cast<long>(cfg.nx) * cfg.ny * cfg.nz;
  for (const BoundaryRun *r : runs) {
    ts << r->ts;
    snap << r->snap;
    const RunSummary &s = r->summary;
    std::cout << cfg.name << '/' << s.material << ": particles=" << particles
              << " runtime_s=" << s.runtimeSeconds << " energy_error=" << s.finalEnergyError
              << " max_eqps=" << s.maxEqps << " pass=" << (s.pass ? "true" : "false") << '\n';
  }
  const bool pass = std::all_of(runs.begin(), runs.end(),
                                [](const BoundaryRun *r) { return r->summary.pass; });
  std::ofstream md(cfg.outputDir + "/metadata.json");
  md << "{\n  \"name\": \"" << cfg.name << "\",\n"
     << "  \"solver\": \"correspondence peridynamics\",\n"
     << "  \"status\": \"" << (pass ? "pass" : "fail") << "\",\n"
     << "  \"classification\": \"" << cfg.classification << "\",\n"
     << "  \"scope\": \"" << cfg.scope << "\",\n"
     << "  \"units\": {\"length\": \"mm\", \"time\": \"ms\", \"force\": \"N\", "
        "\"stress\": \"MPa\", \"solverMassDensity\": \"N*ms^2/mm^4\", "
        "\"sourceDensity\": \"kg/m^3\"},\n"
     << "  \"densityConversion\": \"rho_solver = rho_kg_per_m3 * 1e-6\",\n"
     << "  \"particles\": " << particles << ",\n"
     << "  \"grid\": [" << cfg.nx << ',' << cfg.ny << ',' << cfg.nz
     << "],\n  \"spacing_mm\": " << num(cfg.spacing, 6)
     << ",\n  \"horizon_mm\": " << num(cfg.horizon, 6)
     << ",\n  \"horizonSpacingRatio\": " << num(cfg.horizon / cfg.spacing, 6)
     << ",\n  \"neighborSearch\": \"cell-linked-list\",\n"
     << "  \"planeStrain\": " << (cfg.planeStrain ? "true" : "false")
     << ",\n  \"outerSimpleDamage\": " << (cfg.damageMode != "none" ? "true" : "false") << ",\n"
     << "  \"zeroEnergyStabilization\": \"none\",\n"
     << "  \"timeStepPolicy\": \"user supplied; no computed stability bound\",\n"
     << "  \"boundarySets\": {\"support\": \"" << cfg.supportRegion << "\", \"loaded\": \""
     << cfg.loadRegion << "\", \"regionLayers\": " << cfg.regionLayers
     << ", \"rigidBodySuppression\": \"two lateral components at support centre and z at "
        "support upper midplane point\"},\n"
     << "  \"dt_ms\": " << num(cfg.dt, 6) << ",\n  \"steps\": " << cfg.steps
     << ",\n  \"materials\": [\n";
  for (size_t i = 0; i < runs.size(); ++i) {
    const auto &s = runs[i]->summary;
    const auto &m = *runs[i]->mat;
    md << "    {\"name\": \"" << s.material << "\", \"status\": \"" << (s.pass ? "pass" : "fail")
       << "\", \"plugin_file\": \""
       << std::filesystem::path(m.material.pluginPath).filename().string()
       << "\", \"sourceDensity_kg_per_m3\": " << json_number(m.material.density)
       << ", \"solverDensity\": " << json_number(cfg.density)
       << ", \"runtimeSeconds\": " << json_number(s.runtimeSeconds)
       << ", \"observationAxis\": \"" << cfg.component << "\""
       << ", \"observationMin_mm\": " << json_number(s.observationMin)
       << ", \"observationMax_mm\": " << json_number(s.observationMax)
       << ", \"maxNetInternalForce_N\": " << json_number(s.maxNetInternalForce)
       << ", \"finalEnergyError\": " << json_number(s.finalEnergyError)
       << ", \"maxEnergyError\": " << json_number(s.maxEnergyError)
       << ", \"maxKineticToInternal\": " << json_number(s.maxKineticToInternal)
       << ", \"finalKineticToInternal\": " << json_number(s.finalKineticToInternal)
       << ", \"maxEqps\": " << json_number(s.maxEqps)
       << ", \"maxInternalIndicator\": " << json_number(s.maxInternalIndicator)
       << ", \"earlyElasticSlope_MPa\": " << json_number(s.earlyElasticSlope)
       << ", \"analyticalPlaneStrainSlope_MPa\": " << json_number(s.analyticalPlaneStrainSlope)
       << ", \"earlyElasticRelativeError\": " << json_number(s.earlyElasticRelativeError)
       << ", \"finalReaction_N\": " << json_number(s.finalReaction)
       << ", \"finalNominalStress_MPa\": " << json_number(s.finalNominalStress)
       << ", \"finalEngineeringStrain\": " << json_number(s.finalEngineeringStrain)
       << ", \"finite\": " << (s.finite ? "true" : "false") << "}"
       << (i + 1 == runs.size() ? "\n" : ",\n");
  }
  md << "  ],\n  \"acceptance\": {\"energyToleranceRelative\": " << num(cfg.energyTol, 6)
     << ", \"netForceTolerance_N\": " << num(cfg.netForceTol, 6)
     << ", \"minimumPlasticStrain\": " << num(cfg.minPlastic, 6)
     << ", \"kineticToInternalLimit\": " << num(cfg.kineticLimit, 6) << "},\n"
     << "  \"materialComparisonCaveat\": \"Internal state variables are model-specific observers "
        "and are not mapped to a common bond-breaking rule.\",\n"
     << "  \"parameterMatchingPolicy\": \"Only identifiable macroscopic properties are matched; "
        "formulation-specific or hard-coded parameters are not claimed to be equivalent.\",\n"
     << "  \"legacyBlockPolicy\": \"Material-independent fixed-width blocks of 256 points; a "
        "non-full final block duplicates its last valid point and discards padded outputs.\"\n}\n";
  return pass ? 0 : 2;
}

// runBoundaryValueCase (BoundaryValueCaseRunner.cpp:309-440): the materials one after another.
int run_boundary_case(const CaseConfig &cfg) {
  std::vector<BoundaryRun> runs;
  runs.reserve(cfg.materials.size());
  std::vector<const BoundaryRun *> done;
  for (const auto &m : cfg.materials) {
    runs.emplace_back(cfg, m);
    run_material(runs.back());
    done.push_back(&runs.back());
  }
  return write_boundary_case(cfg, done);
}

// Several boundary cases (every material of every case) as one batch.
int run_boundary_batch(const std::vector<CaseConfig> &cases) {
  std::vector<BoundaryRun> runs;
  size_t total = 0;
  for (const auto &c : cases)
    total += c.materials.size();
  runs.reserve(total);
  for (const auto &c : cases)
    for (const auto &m : c.materials)
      runs.emplace_back(c, m);
  if (runs.empty())
    throw std::runtime_error("No materials to run.");
  run_batch(runs);
  int status = 0;
  size_t k = 0;
  for (const auto &c : cases) {
    std::vector<const BoundaryRun *> mine;
    for (size_t m = 0; m < c.materials.size(); ++m)
      mine.push_back(&runs[k++]);
    status = std::max(status, write_boundary_case(c, mine));
  }
  return status;
}


// ---- affine demo cases (src/host/DemoCaseRunner.cpp) -------------------------------------

// The device VUMAT a material file names (its plugin file or cmname).
int device_material(const CaseMaterial &mat) {
  const std::string plugin = std::filesystem::path(mat.material.pluginPath).filename().string();
  std::string cm = mat.material.cmname;
  std::transform(cm.begin(), cm.end(), cm.begin(), [](unsigned char ch) { return std::toupper(ch); });
  if (cm.rfind("JOHNSONCOOK", 0) == 0 || plugin.rfind("jc_vumat_plugin", 0) == 0)
    return VPD_MATERIAL_JOHNSON_COOK;
  if (cm.rfind("ELASTIC", 0) == 0 || plugin.rfind("elastic_vumat_plugin", 0) == 0)
    return VPD_MATERIAL_ELASTIC;
  throw std::runtime_error("Only the Johnson-Cook and elastic VUMATs run on the device; material " +
                           mat.label + " uses " + plugin);
}

struct DemoSummary { // MaterialRunSummary (DemoCaseRunner.cpp:36-44)
  std::string key;
  bool reference = false;
  double maxFError = 0.0, maxStressReferenceError = 0.0, maxNetForceNorm = 0.0, maxDamage = 0.0;
  bool finite = true;
};

bool all_finite(const double *v, int count) {
  for (int k = 0; k < count; ++k)
    if (!std::isfinite(v[k]))
      return false;
  return true;
}

DemoSummary run_demo_material(const CaseConfig &cfg, const CaseMaterial &mat, std::ostream &ts,
                              std::ostream &field) {
  const int64_t n = static_cast<int64_t>(cfg.nx) * cfg.ny * cfg.nz;
  std::vector<double> x(3 * n), vol(n);
  if (vpd_make_patch(cfg.nx, cfg.ny, cfg.nz, cfg.spacing, x.data(), vol.data()) != VPD_OK)
    throw std::runtime_error("Particle patch requires positive dimensions and spacing.");
  const bool shear = cfg.bcMode == "simple_shear_xy";
  vpd_config vc{};
  vc.horizon = cfg.horizon;
  vc.density = 0.0; // no explicit dynamics in the affine host
  vc.nprops = static_cast<int>(mat.material.props.size());
  vc.props = mat.material.props.data();
  vc.nstatev = mat.material.nstatev;
  vc.initial_state = mat.initialState.empty() ? nullptr : mat.initialState.data();
  vc.reference_temperature = mat.referenceTemperature;
  vc.char_length = cfg.spacing;
  vc.material_density = mat.material.density;
  vc.material_model = device_material(mat);
  if (cfg.damageMode == "critical_stretch") {
    if (cfg.damageInitiation <= 0.0)
      throw std::runtime_error("critical_stretch requires initiation > 0.");
    vc.damage_mode = VPD_DAMAGE_CRITICAL_STRETCH;
  } else if (cfg.damageMode != "none") {
    throw std::runtime_error("Unsupported damage model on the device: " + cfg.damageMode);
  }
  vc.damage_initiation = cfg.damageInitiation;
  vc.damage_failure = cfg.damageFailure;
  vpd_context *raw = nullptr;
  if (vpd_create(&vc, &raw) != VPD_OK)
    throw std::runtime_error(vpd_last_error(nullptr));
  std::unique_ptr<vpd_context, ContextDeleter> owner(raw);
  vpd_context *ctx = raw;
  check(ctx, vpd_set_particles(ctx, x.data(), vol.data(), n));
  int64_t nnz = 0;
  check(ctx, vpd_build_neighbors(ctx, &nnz));
  check(ctx, vpd_initialize_shape(ctx));
  check(ctx, vpd_initialize_material(ctx)); // F at zero displacement (DemoCaseRunner.cpp:325)

  // Elastic references (DemoCaseRunner.cpp:51-62) from the case's elasticE and elasticNu.
  const double lam = mat.elasticE * mat.elasticNu / ((1.0 + mat.elasticNu) * (1.0 - 2.0 * mat.elasticNu));
  const double mu = mat.elasticE / (2.0 * (1.0 + mat.elasticNu));
  double minX = std::numeric_limits<double>::max(), maxX = -std::numeric_limits<double>::max();
  for (int64_t i = 0; i < n; ++i) {
    minX = std::min(minX, x[3 * i]);
    maxX = std::max(maxX, x[3 * i]);
  }
  const int nst = mat.material.nstatev;
  std::vector<double> u(3 * n), F(9 * n), sigma(9 * n), strain(9 * n), force(3 * n),
      dmg(n, 0.0), statev(static_cast<size_t>(std::max(nst, 1)) * n);
  DemoSummary summary;
  summary.key = mat.label;
  summary.reference = mat.reference;
  const int steps = cfg.steps;
  const double dt = cfg.dt;
  for (int step = 1; step <= steps; ++step) {
    // AffineDisplacementBoundaryCondition::apply / loadValue (BoundaryConditions.cpp:24-40)
    const double load = cfg.maxStrain * static_cast<double>(step) / static_cast<double>(steps);
    for (int64_t i = 0; i < n; ++i) {
      u[3 * i] = load * x[3 * i + (shear ? 1 : 0)];
      u[3 * i + 1] = 0.0;
      u[3 * i + 2] = 0.0;
    }
    check(ctx, vpd_upload(ctx, VPD_FIELD_DISPLACEMENT, u.data()));
    check(ctx, vpd_compute_deformation_gradient(ctx));
    check(ctx, vpd_update_material(ctx, dt, static_cast<double>(step - 1) * dt,
                                   static_cast<double>(step - 1) * dt));
    if (vc.damage_mode == VPD_DAMAGE_CRITICAL_STRETCH)
      check(ctx, vpd_update_damage(ctx));
    check(ctx, vpd_download(ctx, VPD_FIELD_STRESS, sigma.data()));
    // Both damage modes leave the stress undegraded: the force state uses the raw stress.
    check(ctx, vpd_compute_force_density(ctx, sigma.data(), force.data()));
    check(ctx, vpd_download(ctx, VPD_FIELD_DEFGRAD, F.data()));
    if (vc.damage_mode == VPD_DAMAGE_CRITICAL_STRETCH)
      check(ctx, vpd_download(ctx, VPD_FIELD_DAMAGE_INDEX, dmg.data()));
    if (nst > 0)
      check(ctx, vpd_download(ctx, VPD_FIELD_STATEV, statev.data()));
    const double refF[9] = {shear ? 1.0 : 1.0 + load, shear ? load : 0.0, 0.0, 0.0, 1.0, 0.0,
                            0.0, 0.0, 1.0};
    const double refXx = shear ? 0.0 : (lam + 2.0 * mu) * load;
    const double refXy = shear ? mu * load : 0.0;
    double maxFError = 0.0;
    for (int64_t i = 0; i < n; ++i)
      for (int c = 0; c < 9; ++c)
        maxFError = std::max(maxFError, std::abs(F[9 * i + c] - refF[c]));
    double rawSum = 0.0, rawXySum = 0.0, damageSum = 0.0, maxDamage = 0.0, maxRefError = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double *s = &sigma[9 * i];
      summary.finite = summary.finite && all_finite(s, 9) && all_finite(&force[3 * i], 3) &&
                       std::isfinite(dmg[i]);
      rawSum += s[0];
      const double sxy = 0.5 * (s[1] + s[3]);
      rawXySum += sxy;
      damageSum += dmg[i];
      maxDamage = std::max(maxDamage, dmg[i]);
      maxRefError = std::max(maxRefError, std::max(std::abs(s[0] - refXx), std::abs(sxy - refXy)));
    }
    const double count = static_cast<double>(n);
    double net[3] = {0.0, 0.0, 0.0}, left = 0.0, right = 0.0; // summarizeForces
    for (int64_t i = 0; i < n; ++i) {
      for (int d = 0; d < 3; ++d)
        net[d] += force[3 * i + d] * vol[i];
      if (std::abs(x[3 * i] - minX) < 1.0e-12)
        left += force[3 * i] * vol[i];
      if (std::abs(x[3 * i] - maxX) < 1.0e-12)
        right += force[3 * i] * vol[i];
    }
    const double netNorm = std::sqrt(net[0] * net[0] + net[1] * net[1] + net[2] * net[2]);
    double indSum = 0.0, indMax = 0.0;
    bool stateOk = true;
    for (int64_t i = 0; i < n; ++i) {
      const double *st = &statev[static_cast<size_t>(nst) * i];
      stateOk = stateOk && all_finite(st, nst);
      const double ind = mat.indicatorIndices.empty() ? 0.0 : std::max(0.0, internal_indicator(mat, st));
      indSum += ind;
      indMax = std::max(indMax, ind);
    }
    const double indAvg = n > 0 ? indSum / count : 0.0;
    summary.maxFError = std::max(summary.maxFError, maxFError);
    summary.maxStressReferenceError = std::max(summary.maxStressReferenceError, maxRefError);
    summary.maxNetForceNorm = std::max(summary.maxNetForceNorm, netNorm);
    summary.maxDamage = std::max(summary.maxDamage, maxDamage);
    summary.finite = summary.finite && all_finite(net, 3) && stateOk && std::isfinite(indAvg) &&
                     std::isfinite(indMax) && std::isfinite(left) && std::isfinite(right);
    const double row[] = {step * dt, load, maxFError, rawSum / count, rawSum / count,
                          rawXySum / count, rawXySum / count, refXx, refXy, maxRefError,
                          maxDamage, damageSum / count, net[0], net[1], net[2], netNorm, left,
                          right, indAvg, indMax};
    ts << mat.label << ',' << step;
    for (double v : row)
      ts << ',' << num(v, 6);
    ts << ',' << (summary.finite ? 1 : 0) << '\n';
  }
  check(ctx, vpd_download(ctx, VPD_FIELD_STRAIN, strain.data()));
  for (int64_t i = 0; i < n; ++i) { // the field_final rows (DemoCaseRunner.cpp:410-426)
    field << mat.label << ',' << i;
    for (int d = 0; d < 3; ++d)
      field << ',' << num(x[3 * i + d], 6);
    for (int d = 0; d < 3; ++d)
      field << ',' << num(u[3 * i + d], 6);
    for (int d = 0; d < 3; ++d)
      field << ',' << num(x[3 * i + d] + u[3 * i + d], 6);
    field << ',';
    for (int c = 0; c < 9; ++c)
      field << num(F[9 * i + c], 6) << ',';
    const double ind =
        mat.indicatorIndices.empty() ? 0.0
                                     : std::max(0.0, internal_indicator(mat, &statev[static_cast<size_t>(nst) * i]));
    field << num(strain[9 * i], 6) << ',' << num(sigma[9 * i], 6) << ',' << num(sigma[9 * i], 6)
          << ',' << num(dmg[i], 6) << ',' << num(ind, 6) << ',' << num(force[3 * i], 6) << ','
          << num(force[3 * i + 1], 6) << ',' << num(force[3 * i + 2], 6) << '\n';
  }
  return summary;
}

std::string json_escape(const std::string &text) {
  std::string out;
  for (char ch : text) {
    switch (ch) {
    case '\\':
      out += "\\\\";
      break;
    case '"':
      out += "\\\"";
      break;
    case '\n':
      out += "\\n";
      break;
    case '\r':
      out += "\\r";
      break;
    case '\t':
      out += "\\t";
      break;
    default:
      out += ch;
    }
  }
  return out;
}

int run_demo_case(const CaseConfig &cfg) {
  std::filesystem::create_directories(cfg.outputDir);
  std::ofstream ts(std::filesystem::path(cfg.outputDir) / "timeseries.csv");
  ts << "material,step,time,applied_strain,max_f_error,avg_raw_stress_xx,"
        "avg_effective_stress_xx,avg_raw_stress_xy,avg_effective_stress_xy,"
        "elastic_reference_stress_xx,elastic_reference_stress_xy,"
        "max_stress_reference_error,max_damage,avg_damage,net_force_x,"
        "net_force_y,net_force_z,net_force_norm,left_reaction_x,"
        "right_reaction_x,avg_internal_indicator,max_internal_indicator,finite\n";
  std::ofstream field(std::filesystem::path(cfg.outputDir) / "field_final.csv");
  field << "material,particle,x0,y0,z0,ux,uy,uz,x,y,z,"
           "Fxx,Fxy,Fxz,Fyx,Fyy,Fyz,Fzx,Fzy,Fzz,"
           "strain_xx,raw_stress_xx,effective_stress_xx,damage,"
           "internal_indicator,force_x,force_y,force_z\n";
  if (cfg.bcType != "affine_displacement")
    throw std::runtime_error("Unsupported boundary condition type: " + cfg.bcType);
  if (cfg.maxStrain == 0.0)
    throw std::runtime_error("Affine displacement boundary condition requires maxStrain.");
  std::vector<DemoSummary> summaries;
  for (const auto &m : cfg.materials)
    summaries.push_back(run_demo_material(cfg, m, ts, field));
  ts.flush();
  field.flush();

  // writeMetadata (DemoCaseRunner.cpp:231-293); the stream keeps precision 16 once set.
  double maxF = 0.0, maxRef = 0.0, maxNet = 0.0, maxDmg = 0.0;
  bool finite = true;
  for (const auto &s : summaries) {
    maxF = std::max(maxF, s.maxFError);
    if (s.reference)
      maxRef = std::max(maxRef, s.maxStressReferenceError);
    maxNet = std::max(maxNet, s.maxNetForceNorm);
    maxDmg = std::max(maxDmg, s.maxDamage);
    finite = finite && s.finite;
  }
  const bool metaPass = finite && maxF < cfg.tolDefgrad &&
                        (summaries.empty() || maxRef < cfg.tolRefStress) && maxNet < cfg.tolNetForce;
  std::ofstream md(std::filesystem::path(cfg.outputDir) / "metadata.json");
  const auto p16 = [](double v) { return num(v, 16); };
  md << "{\n";
  md << "  \"name\": \"" << cfg.name << "\",\n";
  md << "  \"status\": \"" << (metaPass ? "pass" : "fail") << "\",\n";
  md << "  \"scope\": \"" << cfg.scope << "\",\n";
  md << "  \"particles\": " << cfg.nx * cfg.ny * cfg.nz << ",\n";
  md << "  \"steps\": " << cfg.steps << ",\n";
  md << "  \"maxDeformationGradientError\": " << p16(maxF) << ",\n";
  md << "  \"maxReferenceStressError\": " << p16(maxRef) << ",\n";
  md << "  \"maxJcRawStressError\": " << p16(maxRef) << ",\n";
  md << "  \"maxNetForceNorm\": " << p16(maxNet) << ",\n";
  md << "  \"maxDamage\": " << p16(maxDmg) << ",\n";
  md << "  \"allValuesFinite\": " << (finite ? "true" : "false") << ",\n";
  md << "  \"tolerances\": {\"deformationGradientAbsolute\": " << p16(cfg.tolDefgrad)
     << ", \"referenceStressAbsolute\": " << p16(cfg.tolRefStress)
     << ", \"netForceAbsolute\": " << p16(cfg.tolNetForce) << "},\n";
  std::string notes = cfg.scope;
  if (!cfg.checks.empty()) {
    notes += ". Checks: ";
    for (size_t i = 0; i < cfg.checks.size(); ++i)
      notes += (i ? "; " : "") + cfg.checks[i];
  }
  md << "  \"notes\": \"" << json_escape(notes) << "\",\n";
  md << "  \"checks\": [\n";
  for (size_t i = 0; i < cfg.checks.size(); ++i)
    md << "    \"" << cfg.checks[i] << "\"" << (i + 1 == cfg.checks.size() ? "\n" : ",\n");
  md << "  ],\n";
  md << "  \"materials\": [\n";
  for (size_t i = 0; i < summaries.size(); ++i) {
    const auto &s = summaries[i];
    md << "    {\"name\": \"" << s.key << "\", \"reference\": " << (s.reference ? "true" : "false")
       << ", \"maxDeformationGradientError\": " << p16(s.maxFError)
       << ", \"maxStressReferenceError\": " << p16(s.maxStressReferenceError)
       << ", \"maxNetForceNorm\": " << p16(s.maxNetForceNorm)
       << ", \"maxDamage\": " << p16(s.maxDamage)
       << ", \"finite\": " << (s.finite ? "true" : "false") << "}"
       << (i + 1 == summaries.size() ? "\n" : ",\n");
  }
  md << "  ]\n";
  md << "}\n";
  md.close();

  bool pass = true;
  for (const auto &s : summaries) {
    pass = pass && s.finite && s.maxFError < cfg.tolDefgrad && s.maxNetForceNorm < cfg.tolNetForce;
    if (s.reference)
      pass = pass && s.maxStressReferenceError < cfg.tolRefStress;
    std::cout << s.key << ": max_f_error=" << num(s.maxFError, 6)
              << " max_stress_reference_error=" << num(s.maxStressReferenceError, 6)
              << " max_damage=" << num(s.maxDamage, 6)
              << " max_net_force_norm=" << num(s.maxNetForceNorm, 6)
              << " finite=" << (s.finite ? "true" : "false") << "\n";
  }
  std::cout << "VUMAT-PD demo case written to " << cfg.outputDir << "\n";
  return pass ? 0 : 2;
}

} // namespace

extern "C" int vpd_run_cases_json(const char *const *case_paths, int count,
                                  const char *plugin_directory) {
  try {
    if (!case_paths || count < 1)
      throw std::runtime_error("No case files given.");
    std::vector<CaseConfig> cases;
    for (int k = 0; k < count; ++k) {
      if (!case_paths[k])
        throw std::runtime_error("Failed to open JSON file: (null)");
      cases.push_back(load_case(case_paths[k], plugin_directory ? plugin_directory : ""));
      if (cases.back().type != "vumat_pd_boundary_case")
        throw std::runtime_error("Batched runs take vumat_pd_boundary_case files; got: " +
                                 cases.back().type);
    }
    return run_boundary_batch(cases);
  } catch (const std::exception &error) {
    std::cerr << "VUMAT-PD demo case failed: " << error.what() << "\n";
    return 1;
  }
}

extern "C" int vpd_run_case_json(const char *case_path, const char *plugin_directory) {
  try {
    if (!case_path)
      throw std::runtime_error("Failed to open JSON file: (null)");
    const CaseConfig cfg = load_case(case_path, plugin_directory ? plugin_directory : "");
    if (cfg.type == "vumat_pd_demo_case")
      return run_demo_case(cfg);
    if (cfg.type == "vumat_material_point_case")
      throw std::runtime_error("The GPU case runner runs vumat_pd_boundary_case and "
                               "vumat_pd_demo_case files; got: vumat_material_point_case");
    return run_boundary_case(cfg);
  } catch (const std::exception &error) {
    std::cerr << "VUMAT-PD demo case failed: " << error.what() << "\n";
    return 1;
  }
}
