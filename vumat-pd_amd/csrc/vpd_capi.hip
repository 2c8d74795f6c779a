// C ABI of the MI355X PD increment (include/vpd.h): context lifetime, device buffers,
// kernel launches on the context stream, state transfer and error mapping.
#include "vpd.h"
#include "vpd_kernels.hpp"
#include "vpd_neighbors.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

using namespace vpd;

namespace {

thread_local std::string g_error;

enum KernelId {
  K_PREDICT,
  K_CONSTRAIN,
  K_DEFGRAD,
  K_MATERIAL,
  K_FORCE,
  K_SHAPE,
  K_NEIGHBORS,
  K_COUNT
};
const char *kKernelNames[K_COUNT] = {"predict",      "constrain", "defgrad_bond_pass",
                                     "material_jc", "force_bond_pass", "shape_tensor",
                                     "neighbor_build"};

struct TimingRec {
  int kernel;
  hipEvent_t start, stop;
};

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace

struct vpd_context {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // configuration
  double horizon = 0, density = 0, charLength = 1, tref = 293, materialDensity = 0;
  std::vector<double> props, initialState;
  int nstatev = 0, damageMode = 0, returnMap = 0;
  double initiation = 0, failure = 1;
  JcParams jc{};
  // block
  int64_t n = 0, nnz = 0;
  double *x = nullptr, *vol = nullptr;
  int *rowptr = nullptr, *col = nullptr;
  uint8_t *codes = nullptr;
  double *kinv = nullptr;
  double *u = nullptr, *v = nullptr, *a = nullptr;
  double *fN = nullptr, *vN = nullptr, *rN = nullptr, *sig = nullptr, *statev = nullptr;
  double *temp = nullptr, *eint = nullptr, *einel = nullptr;
  double *fnew = nullptr, *amap = nullptr;
  double *sigOut = nullptr, *strainOut = nullptr, *force = nullptr, *dmgIndex = nullptr;
  int *iters = nullptr;
  uint32_t *broken = nullptr; // one bit per sliced bond slot
  int64_t brokenWords = 0;
  int *colS = nullptr;         // sliced bond layout (see Sell in vpd_kernels.hpp)
  int64_t *sliceBase = nullptr;
  int64_t sellTotal = 0;
  std::vector<int> hostRowptr;
  std::vector<int64_t> hostBase;
  int *derr = nullptr;
  bool haveNeighbors = false, haveShape = false, haveMaterial = false, haveDynamics = false;
  bool haveFnew = false, haveFields = false, haveForce = false, predicted = false;
  bool timing = false;
  std::vector<TimingRec> recs;
  std::vector<hipEvent_t> eventPool;
};

namespace {

int fail(vpd_context *c, int status, const std::string &msg) {
  if (c)
    c->err = msg;
  g_error = msg;
  return status;
}

#define HIPTRY(c, expr)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return fail((c), VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_) +    \
                                           " at " #expr);                                      \
  } while (0)

template <class T> int dalloc(vpd_context *c, T *&p, int64_t count) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  if (count <= 0)
    count = 1;
  HIPTRY(c, hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * (size_t)count));
  return VPD_OK;
}

template <class T> void dfree(T *&p) {
  if (p)
    (void)hipFree(p);
  p = nullptr;
}

hipEvent_t take_event(vpd_context *c) {
  if (!c->eventPool.empty()) {
    hipEvent_t e = c->eventPool.back();
    c->eventPool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Runs `launch` on the context stream, bracketed by events when timing is on.
template <class F> int timed(vpd_context *c, int kernel, F &&launch) {
  TimingRec r{kernel, nullptr, nullptr};
  if (c->timing) {
    r.start = take_event(c);
    r.stop = take_event(c);
    HIPTRY(c, hipEventRecord(r.start, c->stream));
  }
  launch();
  HIPTRY(c, hipGetLastError());
  if (c->timing) {
    HIPTRY(c, hipEventRecord(r.stop, c->stream));
    c->recs.push_back(r);
  }
  return VPD_OK;
}

JcParams make_jc(const std::vector<double> &p) {
  // JohnsonCook.f90:55-72
  JcParams j{};
  j.young = p[0];
  j.poisson = p[1];
  j.a0 = p[2];
  j.b0 = p[3];
  j.rateC = p[4];
  j.hardN = p[5];
  j.thermalM = p[6];
  j.rate0 = p[7];
  j.tref = p[8];
  j.tmelt = p[9];
  j.rho = p[10];
  j.beta = p[11];
  j.cp = p[12];
  j.workToHeat = p[13];
  j.mu = j.young / (2.0 * (1.0 + j.poisson));
  j.lambda = j.young * j.poisson / ((1.0 + j.poisson) * (1.0 - 2.0 * j.poisson));
  j.tolerance = j.young * 1.0e-9;
  return j;
}

// Maps the device error bits to the reference's exception messages.
// Reports (and clears) the error bits in `mask`; other bits stay pending for a later check.
// The material update leaves the first-Piola bit pending: the reference raises it only when
// the force state is assembled (PdCorrespondenceSkeleton.cpp:83-91, called from :209).
int check_device(vpd_context *c, int mask = ~0) {
  int all = 0;
  HIPTRY(c, hipMemcpyAsync(&all, c->derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  const int bits = all & mask;
  if (!bits)
    return VPD_OK;
  const int keep = all & ~mask;
  HIPTRY(c, hipMemcpyAsync(c->derr, &keep, sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  if (bits & ERR_NEIGHBOR_RANGE)
    return fail(c, VPD_ERR_ARGUMENT, "PD neighbor index is out of range.");
  if (bits & ERR_NONFINITE_COORD)
    return fail(c, VPD_ERR_NUMERIC, "Nonfinite PD coordinate.");
  if (bits & ERR_SINGULAR_SHAPE)
    return fail(c, VPD_ERR_NUMERIC, "Singular tensor in PD shape tensor.");
  if (bits & ERR_SINGULAR_COROT)
    return fail(c, VPD_ERR_NUMERIC, "Corotational update encountered a singular tensor.");
  if (bits & ERR_PIOLA_DET)
    return fail(c, VPD_ERR_NUMERIC, "Invalid deformation gradient determinant.");
  return fail(c, VPD_ERR_NUMERIC, "Device reported error bits " + std::to_string(bits));
}

Planes planes(double *p, int64_t n) { return Planes{p, n}; }

int grid_of(int64_t n) { return (int)ceil_div(n, 256); }
// Grids of the XCD-remapped kernels are rounded up to whole rounds of 8 workgroups.
int grid8_of(int64_t n) { return (int)(ceil_div(ceil_div(n, 256), 8) * 8); }

int components(const vpd_context *c, int field) {
  switch (field) {
  case VPD_FIELD_REFERENCE_POSITION:
  case VPD_FIELD_DISPLACEMENT:
  case VPD_FIELD_VELOCITY:
  case VPD_FIELD_ACCELERATION:
  case VPD_FIELD_FORCE_DENSITY:
    return 3;
  case VPD_FIELD_VOLUME:
  case VPD_FIELD_TEMPERATURE:
  case VPD_FIELD_INTERNAL_ENERGY:
  case VPD_FIELD_INELASTIC_ENERGY:
  case VPD_FIELD_DAMAGE_INDEX:
    return 1;
  case VPD_FIELD_STATEV:
    return c->nstatev;
  default:
    return 9;
  }
}

double *field_ptr(vpd_context *c, int field) {
  switch (field) {
  case VPD_FIELD_REFERENCE_POSITION:
    return c->x;
  case VPD_FIELD_VOLUME:
    return c->vol;
  case VPD_FIELD_DISPLACEMENT:
    return c->u;
  case VPD_FIELD_VELOCITY:
    return c->v;
  case VPD_FIELD_ACCELERATION:
    return c->a;
  case VPD_FIELD_DEFGRAD_N:
    return c->fN;
  case VPD_FIELD_LEFT_STRETCH_N:
    return c->vN;
  case VPD_FIELD_ROTATION_N:
    return c->rN;
  case VPD_FIELD_STATEV:
    return c->statev;
  case VPD_FIELD_TEMPERATURE:
    return c->temp;
  case VPD_FIELD_INTERNAL_ENERGY:
    return c->eint;
  case VPD_FIELD_INELASTIC_ENERGY:
    return c->einel;
  case VPD_FIELD_STRESS:
    return c->sigOut;
  case VPD_FIELD_STRAIN:
    return c->strainOut;
  case VPD_FIELD_FORCE_DENSITY:
    return c->force;
  case VPD_FIELD_DAMAGE_INDEX:
    return c->dmgIndex;
  case VPD_FIELD_SHAPE_INVERSE:
    return c->kinv;
  case VPD_FIELD_DEFGRAD:
    return c->fnew;
  default:
    return nullptr;
  }
}

// Host AoS rows <-> device SoA planes.
int put_planes(vpd_context *c, double *dst, const double *host, int comps) {
  std::vector<double> stage((size_t)c->n * comps);
  for (int64_t i = 0; i < c->n; ++i)
    for (int k = 0; k < comps; ++k)
      stage[(size_t)k * c->n + i] = host[(size_t)i * comps + k];
  HIPTRY(c, hipMemcpyAsync(dst, stage.data(), sizeof(double) * stage.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int get_planes(vpd_context *c, const double *src, double *host, int comps) {
  std::vector<double> stage((size_t)c->n * comps);
  HIPTRY(c, hipMemcpyAsync(stage.data(), src, sizeof(double) * stage.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < c->n; ++i)
    for (int k = 0; k < comps; ++k)
      host[(size_t)i * comps + k] = stage[(size_t)k * c->n + i];
  return VPD_OK;
}

__global__ void k_material_init(int64_t n, Planes vN, Planes rN, Planes sig, Planes statev,
                                int nstatev, const double *__restrict__ init, double *temp,
                                double tref, double *eint, double *einel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  for (int c = 0; c < 9; ++c) {
    const double e = (c % 4 == 0) ? 1.0 : 0.0;
    vN(c, i) = e;
    rN(c, i) = e;
  }
  for (int c = 0; c < 6; ++c)
    sig(c, i) = 0.0;
  for (int c = 0; c < nstatev; ++c)
    statev(c, i) = init[c];
  temp[i] = tref;
  eint[i] = 0.0;
  einel[i] = 0.0;
}

// Exclusive scan of `count` ints into int32 `out`; *total receives the sum.
int scan_ints(vpd_context *c, const int *in, int *out, int64_t count, int64_t *total) {
  const int64_t nb = ceil_div(count, SCAN_ITEMS);
  int64_t *part = nullptr, *bt = nullptr;
  if (dalloc(c, part, count) || dalloc(c, bt, nb + 1))
    return VPD_ERR_DEVICE;
  k_scan_block<<<(int)nb, 256, 0, c->stream>>>(count, in, part, bt);
  k_scan_totals<<<1, 256, 0, c->stream>>>(nb, bt, bt + nb);
  HIPTRY(c, hipMemcpyAsync(total, bt + nb, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  if (*total > INT_MAX) {
    dfree(part);
    dfree(bt);
    return fail(c, VPD_ERR_UNSUPPORTED, "Neighbour count exceeds the int32 CSR range.");
  }
  k_scan_add<<<grid_of(count), 256, 0, c->stream>>>(count, part, bt, out);
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  dfree(part);
  dfree(bt);
  return VPD_OK;
}

int ensure_outputs(vpd_context *c) {
  if (!c->sigOut && dalloc(c, c->sigOut, 9 * c->n))
    return VPD_ERR_DEVICE;
  if (!c->strainOut && dalloc(c, c->strainOut, 9 * c->n))
    return VPD_ERR_DEVICE;
  if (!c->force && dalloc(c, c->force, 3 * c->n))
    return VPD_ERR_DEVICE;
  if (!c->iters && dalloc(c, c->iters, c->n))
    return VPD_ERR_DEVICE;
  return VPD_OK;
}

// Sliced copy of the bond lists: slice s holds particles 64s..64s+63, padded to its
// longest row; bond flags are indexed by sliced slot.
int build_sell(vpd_context *c) {
  const int64_t n = c->n;
  c->hostRowptr.resize(n + 1);
  HIPTRY(c, hipMemcpyAsync(c->hostRowptr.data(), c->rowptr, sizeof(int) * (n + 1),
                           hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  const int64_t ns = ceil_div(n, 64);
  c->hostBase.assign(ns + 1, 0);
  for (int64_t s = 0; s < ns; ++s) {
    int m = 0;
    for (int64_t i = 64 * s; i < std::min<int64_t>(n, 64 * s + 64); ++i)
      m = std::max(m, c->hostRowptr[i + 1] - c->hostRowptr[i]);
    c->hostBase[s + 1] = c->hostBase[s] + 64 * (int64_t)m;
  }
  c->sellTotal = c->hostBase[ns];
  if (dalloc(c, c->sliceBase, ns + 1) || dalloc(c, c->colS, c->sellTotal))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemcpyAsync(c->sliceBase, c->hostBase.data(), sizeof(int64_t) * (ns + 1),
                           hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipMemsetAsync(c->colS, 0, sizeof(int) * std::max<int64_t>(1, c->sellTotal),
                           c->stream));
  k_sell_fill<<<grid_of(n), 256, 0, c->stream>>>(n, c->rowptr, c->col, c->sliceBase, c->colS);
  HIPTRY(c, hipGetLastError());
  c->brokenWords = ceil_div(c->sellTotal, 32);
  if (dalloc(c, c->broken, c->brokenWords))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemsetAsync(c->broken, 0, sizeof(uint32_t) * std::max<int64_t>(1, c->brokenWords),
                           c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

// Sliced slot of CSR entry p of row i.
int64_t sell_slot(const vpd_context *c, int64_t i, int64_t p) {
  return c->hostBase[i >> 6] + 64 * (p - c->hostRowptr[i]) + (i & 63);
}

Sell sell_of(const vpd_context *c) { return Sell{c->sliceBase, c->colS, c->rowptr}; }

Ramp ramp_of(const double r[3]) { return Ramp{r[0], r[1], r[2]}; }

int launch_predict(vpd_context *c, double dt, const double ramp[3]) {
  return timed(c, K_PREDICT, [&] {
    k_predict<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, planes(c->u, c->n), planes(c->v, c->n),
                                                    planes(c->a, c->n), c->codes, dt,
                                                    ramp_of(ramp));
  });
}

int launch_defgrad(vpd_context *c) {
  const int64_t nblocks = ceil_div(c->n, 256);
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  const int useCompression = c->failure > 0.0 && std::isfinite(c->failure);
  return timed(c, K_DEFGRAD, [&] {
    if (dmg)
      k_defgrad<true><<<grid8_of(c->n), 256, 0, c->stream>>>(
          c->n, nblocks, planes(c->x, c->n), planes(c->u, c->n), c->vol, sell_of(c),
          planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->initiation, c->failure,
          useCompression);
    else
      k_defgrad<false><<<grid8_of(c->n), 256, 0, c->stream>>>(
          c->n, nblocks, planes(c->x, c->n), planes(c->u, c->n), c->vol, sell_of(c),
          planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->initiation, c->failure,
          useCompression);
  });
}

int launch_material(vpd_context *c, double dt, bool fields) {
  MaterialArgs g;
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.fnew = planes(c->fnew, c->n);
  g.fN = planes(c->fN, c->n);
  g.vN = planes(c->vN, c->n);
  g.rN = planes(c->rN, c->n);
  g.sig = planes(c->sig, c->n);
  g.statev = planes(c->statev, c->n);
  g.kinv = planes(c->kinv, c->n);
  g.amap = planes(c->amap, c->n);
  g.temp = c->temp;
  g.eint = c->eint;
  g.einel = c->einel;
  g.sigOut = planes(fields ? c->sigOut : nullptr, c->n);
  g.strainOut = planes(fields ? c->strainOut : nullptr, c->n);
  g.iterOut = fields ? c->iters : nullptr;
  g.jc = c->jc;
  g.dt = dt;
  g.err = c->derr;
  return timed(c, K_MATERIAL, [&] { k_material<<<grid8_of(c->n), 256, 0, c->stream>>>(g); });
}

int launch_force(vpd_context *c, double dt, const double ramp[3], bool fuse,
                 const double *rampNext, bool fields) {
  ForceArgs g;
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.x = planes(c->x, c->n);
  g.u = planes(c->u, c->n);
  g.v = planes(c->v, c->n);
  g.a = planes(c->a, c->n);
  g.amap = planes(c->amap, c->n);
  g.forceOut = planes(fields ? c->force : nullptr, c->n);
  g.vol = c->vol;
  g.bonds = sell_of(c);
  g.broken = c->broken;
  g.codes = c->codes;
  g.dt = dt;
  g.density = c->density;
  g.rampNow = ramp_of(ramp);
  g.rampNext = fuse ? ramp_of(rampNext) : Ramp{0, 0, 0};
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  const int grid = grid8_of(c->n);
  return timed(c, K_FORCE, [&] {
    if (dmg && fuse)
      k_force<true, true><<<grid, 256, 0, c->stream>>>(g);
    else if (dmg)
      k_force<true, false><<<grid, 256, 0, c->stream>>>(g);
    else if (fuse)
      k_force<false, true><<<grid, 256, 0, c->stream>>>(g);
    else
      k_force<false, false><<<grid, 256, 0, c->stream>>>(g);
  });
}

int require_block(vpd_context *c) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  if (c->n <= 0)
    return fail(c, VPD_ERR_STATE, "Invalid or incomplete PD particle block.");
  return VPD_OK;
}

int require_ready(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (!c->haveNeighbors || !c->haveShape)
    return fail(c, VPD_ERR_STATE, "PD correspondence state is not initialized.");
  if (!c->haveMaterial)
    return fail(c, VPD_ERR_STATE,
                "VUMAT material state is not initialized for this point block.");
  if (!c->haveDynamics)
    return fail(c, VPD_ERR_STATE, "Invalid explicit dynamics predictor input.");
  return VPD_OK;
}

} // namespace

extern "C" {

int vpd_abi_version(void) { return VPD_ABI_VERSION; }

int vpd_device_count(int *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (count)
    *count = e == hipSuccess ? n : 0;
  return e == hipSuccess ? VPD_OK : fail(nullptr, VPD_ERR_DEVICE, "No HIP device available.");
}

const char *vpd_last_error(const vpd_context *ctx) {
  return ctx ? ctx->err.c_str() : g_error.c_str();
}

int vpd_create(const vpd_config *cfg, vpd_context **out) {
  if (!cfg || !out)
    return fail(nullptr, VPD_ERR_ARGUMENT, "vpd_create requires a config and an output.");
  *out = nullptr;
  if (!(cfg->horizon > 0.0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "PD correspondence skeleton requires horizon > 0.");
  if (cfg->density < 0.0 || std::isnan(cfg->density))
    return fail(nullptr, VPD_ERR_ARGUMENT, "Explicit dynamics requires density > 0.");
  const bool material = cfg->nprops > 0 || cfg->props;
  if (material && !(cfg->char_length > 0.0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "VUMAT-PD material bridge requires charLength > 0.");
  if (material && (cfg->nprops <= 0 || !cfg->props))
    return fail(nullptr, VPD_ERR_ARGUMENT, "VUMAT props cannot be empty.");
  if (material && (cfg->nprops < 14 || cfg->nstatev < 6))
    return fail(nullptr, VPD_ERR_ARGUMENT, "JohnsonCook VUMAT requires 14 props and 6 states");
  if (cfg->damage_mode != VPD_DAMAGE_NONE && cfg->damage_mode != VPD_DAMAGE_CRITICAL_STRETCH)
    return fail(nullptr, VPD_ERR_UNSUPPORTED, "Unsupported damage mode.");
  if (cfg->damage_mode == VPD_DAMAGE_CRITICAL_STRETCH && !(cfg->damage_initiation > 0.0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "critical_stretch requires initiation > 0.");
  if (cfg->return_map != VPD_RETURN_MAP_BISECTION)
    return fail(nullptr, VPD_ERR_UNSUPPORTED, "Unsupported return-map mode.");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, VPD_ERR_DEVICE, "No HIP device available.");
  auto *c = new vpd_context();
  c->horizon = cfg->horizon;
  c->density = cfg->density;
  c->charLength = cfg->char_length;
  c->tref = cfg->reference_temperature;
  c->materialDensity = cfg->material_density;
  if (material)
    c->props.assign(cfg->props, cfg->props + cfg->nprops);
  c->nstatev = material ? cfg->nstatev : 0;
  c->initialState.assign(c->nstatev, 0.0);
  if (material && cfg->initial_state)
    c->initialState.assign(cfg->initial_state, cfg->initial_state + cfg->nstatev);
  c->damageMode = cfg->damage_mode;
  c->initiation = cfg->damage_initiation;
  c->failure = cfg->damage_failure;
  c->returnMap = cfg->return_map;
  if (material)
    c->jc = make_jc(c->props);
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&c->derr), sizeof(int)) != hipSuccess ||
      hipMemset(c->derr, 0, sizeof(int)) != hipSuccess) {
    vpd_destroy(c);
    return fail(nullptr, VPD_ERR_DEVICE, "HIP stream or allocation failed.");
  }
  *out = c;
  return VPD_OK;
}

void vpd_destroy(vpd_context *c) {
  if (!c)
    return;
  if (c->stream)
    (void)hipStreamSynchronize(c->stream);
  for (auto &r : c->recs) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : c->eventPool)
    (void)hipEventDestroy(e);
  dfree(c->x);
  dfree(c->vol);
  dfree(c->rowptr);
  dfree(c->col);
  dfree(c->codes);
  dfree(c->kinv);
  dfree(c->u);
  dfree(c->v);
  dfree(c->a);
  dfree(c->fN);
  dfree(c->vN);
  dfree(c->rN);
  dfree(c->sig);
  dfree(c->statev);
  dfree(c->temp);
  dfree(c->eint);
  dfree(c->einel);
  dfree(c->fnew);
  dfree(c->amap);
  dfree(c->sigOut);
  dfree(c->strainOut);
  dfree(c->force);
  dfree(c->dmgIndex);
  dfree(c->iters);
  dfree(c->broken);
  dfree(c->colS);
  dfree(c->sliceBase);
  dfree(c->derr);
  if (c->stream)
    (void)hipStreamDestroy(c->stream);
  delete c;
}

int vpd_set_particles(vpd_context *c, const double *x, const double *vol, int64_t n) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  if (n <= 0 || !x || !vol)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid or incomplete PD particle block.");
  if (n > INT_MAX)
    return fail(c, VPD_ERR_UNSUPPORTED, "Particle count exceeds the int32 index range.");
  c->n = n;
  const int64_t sz[] = {3 * n, n, 9 * n, 3 * n, 3 * n, 3 * n, 9 * n, 9 * n,
                        9 * n, 6 * n, (int64_t)c->nstatev * n, n, n, n, 9 * n, 9 * n};
  double **bufs[] = {&c->x,  &c->vol, &c->kinv, &c->u,    &c->v,   &c->a,
                     &c->fN, &c->vN,  &c->rN,   &c->sig,  &c->statev, &c->temp,
                     &c->eint, &c->einel, &c->fnew, &c->amap};
  for (size_t k = 0; k < sizeof(sz) / sizeof(sz[0]); ++k)
    if (dalloc(c, *bufs[k], sz[k]))
      return VPD_ERR_DEVICE;
  if (dalloc(c, c->codes, n))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemsetAsync(c->codes, 0, n, c->stream));
  HIPTRY(c, hipMemsetAsync(c->u, 0, sizeof(double) * 3 * n, c->stream));
  dfree(c->sigOut);
  dfree(c->strainOut);
  dfree(c->force);
  dfree(c->dmgIndex);
  dfree(c->iters);
  c->haveNeighbors = c->haveShape = c->haveMaterial = c->haveDynamics = false;
  c->haveFnew = c->haveFields = c->haveForce = c->predicted = false;
  if (int s = put_planes(c, c->x, x, 3))
    return s;
  HIPTRY(c, hipMemcpyAsync(c->vol, vol, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int vpd_set_neighbors(vpd_context *c, const int32_t *rowptr, const int32_t *col, int64_t nnz) {
  if (int s = require_block(c))
    return s;
  if (!rowptr || (nnz > 0 && !col) || nnz < 0 || nnz > INT_MAX)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid or incomplete PD particle block.");
  if (rowptr[0] != 0 || rowptr[c->n] != nnz)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid or incomplete PD particle block.");
  for (int64_t i = 0; i < c->n; ++i)
    if (rowptr[i + 1] < rowptr[i])
      return fail(c, VPD_ERR_ARGUMENT, "Invalid or incomplete PD particle block.");
  for (int64_t p = 0; p < nnz; ++p)
    if (col[p] < 0 || col[p] >= c->n)
      return fail(c, VPD_ERR_ARGUMENT, "PD neighbor index is out of range.");
  c->nnz = nnz;
  if (dalloc(c, c->rowptr, c->n + 1) || dalloc(c, c->col, nnz))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemcpyAsync(c->rowptr, rowptr, sizeof(int) * (c->n + 1), hipMemcpyHostToDevice,
                           c->stream));
  if (nnz)
    HIPTRY(c, hipMemcpyAsync(c->col, col, sizeof(int) * nnz, hipMemcpyHostToDevice, c->stream));
  if (int st = build_sell(c))
    return st;
  c->haveNeighbors = true;
  c->haveShape = c->haveMaterial = false;
  return VPD_OK;
}

int vpd_build_neighbors(vpd_context *c, int64_t *nnzOut) {
  if (int s = require_block(c))
    return s;
  const int64_t n = c->n;
  Bounds hb;
  hb.maxAbsBits = 0;
  for (int d = 0; d < 3; ++d) {
    hb.cmin[d] = LLONG_MAX;
    hb.cmax[d] = LLONG_MIN;
  }
  hb.nonfinite = 0;
  Bounds *db = nullptr;
  if (dalloc(c, db, 1))
    return VPD_ERR_DEVICE;
  auto cleanup = [&] { dfree(db); };
  HIPTRY(c, hipMemcpyAsync(db, &hb, sizeof(Bounds), hipMemcpyHostToDevice, c->stream));
  TimingRec rec{K_NEIGHBORS, nullptr, nullptr};
  if (c->timing) {
    rec.start = take_event(c);
    rec.stop = take_event(c);
    HIPTRY(c, hipEventRecord(rec.start, c->stream));
  }
  k_abs_max<<<grid_of(n), 256, 0, c->stream>>>(n, c->x, db);
  HIPTRY(c, hipMemcpyAsync(&hb, db, sizeof(Bounds), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  if (hb.nonfinite) {
    cleanup();
    return fail(c, VPD_ERR_NUMERIC, "Nonfinite PD coordinate.");
  }
  double maxAbs;
  std::memcpy(&maxAbs, &hb.maxAbsBits, sizeof(double));
  // PdCorrespondenceSkeleton.cpp:30-39, 101-102
  const double scale = std::max(c->horizon, maxAbs);
  const double tol = 64.0 * std::numeric_limits<double>::epsilon() * scale;
  Grid g;
  g.radius = c->horizon + tol;
  g.cw = c->horizon + 2.0 * tol;
  k_cell_bounds<<<grid_of(n), 256, 0, c->stream>>>(n, c->x, g.cw, db);
  HIPTRY(c, hipMemcpyAsync(&hb, db, sizeof(Bounds), hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  cleanup();
  double cells = 1.0;
  for (int d = 0; d < 3; ++d) {
    g.c0[d] = hb.cmin[d];
    g.dim[d] = hb.cmax[d] - hb.cmin[d] + 1;
    cells *= (double)g.dim[d];
  }
  if (cells > (double)(1 << 28))
    return fail(c, VPD_ERR_UNSUPPORTED,
                "Cell grid too large for the GPU neighbour search (more than 2^28 cells).");
  const int64_t nc = (int64_t)cells;
  int *cellOf = nullptr, *cellCount = nullptr, *cellStart = nullptr, *cursor = nullptr,
      *items = nullptr, *rowCount = nullptr;
  if (dalloc(c, cellOf, n) || dalloc(c, cellCount, nc) || dalloc(c, cellStart, nc) ||
      dalloc(c, cursor, nc) || dalloc(c, items, n) || dalloc(c, rowCount, n + 1))
    return VPD_ERR_DEVICE;
  auto freeAll = [&] {
    dfree(cellOf);
    dfree(cellCount);
    dfree(cellStart);
    dfree(cursor);
    dfree(items);
    dfree(rowCount);
  };
  HIPTRY(c, hipMemsetAsync(cellCount, 0, sizeof(int) * nc, c->stream));
  HIPTRY(c, hipMemsetAsync(cursor, 0, sizeof(int) * nc, c->stream));
  HIPTRY(c, hipMemsetAsync(rowCount, 0, sizeof(int) * (n + 1), c->stream));
  k_cell_count<<<grid_of(n), 256, 0, c->stream>>>(n, c->x, g, cellOf, cellCount);
  int64_t placed = 0;
  if (int s = scan_ints(c, cellCount, cellStart, nc, &placed)) {
    freeAll();
    return s;
  }
  k_cell_fill<<<grid_of(n), 256, 0, c->stream>>>(n, cellOf, cellStart, cursor, items);
  if (dalloc(c, c->rowptr, n + 1)) {
    freeAll();
    return VPD_ERR_DEVICE;
  }
  k_row_scan<false><<<grid_of(n), 256, 0, c->stream>>>(n, c->x, g, cellStart, cellCount, items,
                                                       rowCount, nullptr, nullptr);
  int64_t nnz = 0;
  if (int s = scan_ints(c, rowCount, c->rowptr, n + 1, &nnz)) {
    freeAll();
    return s;
  }
  if (dalloc(c, c->col, nnz)) {
    freeAll();
    return VPD_ERR_DEVICE;
  }
  k_row_scan<true><<<grid_of(n), 256, 0, c->stream>>>(n, c->x, g, cellStart, cellCount, items,
                                                      nullptr, c->rowptr, c->col);
  k_row_sort<<<(int)n, 64, 0, c->stream>>>(n, c->rowptr, c->col);
  HIPTRY(c, hipGetLastError());
  if (c->timing) {
    HIPTRY(c, hipEventRecord(rec.stop, c->stream));
    c->recs.push_back(rec);
  }
  HIPTRY(c, hipStreamSynchronize(c->stream));
  freeAll();
  c->nnz = nnz;
  if (int st = build_sell(c))
    return st;
  c->haveNeighbors = true;
  c->haveShape = c->haveMaterial = false;
  if (nnzOut)
    *nnzOut = nnz;
  return VPD_OK;
}

int vpd_get_neighbors(vpd_context *c, int32_t *rowptr, int32_t *col) {
  if (int s = require_block(c))
    return s;
  if (!c->haveNeighbors)
    return fail(c, VPD_ERR_STATE, "Neighbour lists have not been built.");
  if (rowptr)
    HIPTRY(c, hipMemcpyAsync(rowptr, c->rowptr, sizeof(int) * (c->n + 1), hipMemcpyDeviceToHost,
                             c->stream));
  if (col && c->nnz)
    HIPTRY(c, hipMemcpyAsync(col, c->col, sizeof(int) * c->nnz, hipMemcpyDeviceToHost,
                             c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int vpd_set_constraints(vpd_context *c, const uint8_t *codes) {
  if (int s = require_block(c))
    return s;
  if (!codes) {
    HIPTRY(c, hipMemsetAsync(c->codes, 0, c->n, c->stream));
  } else {
    for (int64_t i = 0; i < c->n; ++i)
      for (int d = 0; d < 3; ++d)
        if (((codes[i] >> (2 * d)) & 3) == 3)
          return fail(c, VPD_ERR_ARGUMENT, "Invalid constraint code.");
    HIPTRY(c, hipMemcpyAsync(c->codes, codes, c->n, hipMemcpyHostToDevice, c->stream));
  }
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int vpd_initialize_shape(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (!c->haveNeighbors)
    return fail(c, VPD_ERR_STATE, "Neighbour lists have not been built.");
  if (int s = timed(c, K_SHAPE, [&] {
        k_shape<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, planes(c->x, c->n), c->vol,
                                                      c->rowptr, c->col, planes(c->kinv, c->n),
                                                      c->derr);
      }))
    return s;
  if (int s = check_device(c))
    return s;
  c->haveShape = true;
  return VPD_OK;
}

int vpd_initialize_material(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (c->props.empty())
    return fail(c, VPD_ERR_STATE, "VUMAT props cannot be empty.");
  if (!c->haveShape)
    return fail(c, VPD_ERR_STATE, "PD correspondence state is not initialized.");
  // F₀ at the current displacement into fN (VumatPdCorrespondenceMaterial.cpp:54).
  const int saved = c->damageMode;
  c->damageMode = VPD_DAMAGE_NONE;
  int s = launch_defgrad(c);
  c->damageMode = saved;
  if (s)
    return s;
  HIPTRY(c, hipMemcpyAsync(c->fN, c->fnew, sizeof(double) * 9 * c->n, hipMemcpyDeviceToDevice,
                           c->stream));
  double *init = nullptr;
  if (dalloc(c, init, c->nstatev))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemcpyAsync(init, c->initialState.data(), sizeof(double) * c->nstatev,
                           hipMemcpyHostToDevice, c->stream));
  k_material_init<<<grid_of(c->n), 256, 0, c->stream>>>(
      c->n, planes(c->vN, c->n), planes(c->rN, c->n), planes(c->sig, c->n),
      planes(c->statev, c->n), c->nstatev, init, c->temp, c->tref, c->eint, c->einel);
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  dfree(init);
  c->haveMaterial = true;
  c->haveFnew = true;
  return VPD_OK;
}

int vpd_initialize_dynamics(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (!(c->density > 0.0))
    return fail(c, VPD_ERR_STATE, "Explicit dynamics requires density > 0.");
  HIPTRY(c, hipMemsetAsync(c->v, 0, sizeof(double) * 3 * c->n, c->stream));
  HIPTRY(c, hipMemsetAsync(c->a, 0, sizeof(double) * 3 * c->n, c->stream));
  if (c->broken)
    HIPTRY(c, hipMemsetAsync(c->broken, 0,
                             sizeof(uint32_t) * std::max<int64_t>(1, c->brokenWords), c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->haveDynamics = true;
  c->predicted = false;
  return VPD_OK;
}

int vpd_upload(vpd_context *c, int field, const void *host) {
  if (int s = require_block(c))
    return s;
  if (!host)
    return fail(c, VPD_ERR_ARGUMENT, "Null host buffer.");
  if (field == VPD_FIELD_BOND_DAMAGE) {
    if (!c->haveNeighbors)
      return fail(c, VPD_ERR_STATE, "Neighbour lists have not been built.");
    const double *d = static_cast<const double *>(host);
    std::vector<uint32_t> bits(std::max<int64_t>(1, c->brokenWords), 0u);
    for (int64_t i = 0; i < c->n; ++i)
      for (int64_t p = c->hostRowptr[i]; p < c->hostRowptr[i + 1]; ++p) {
        const int64_t q = sell_slot(c, i, p);
        if (std::min(std::max(d[p], 0.0), 1.0) >= 1.0)
          bits[q >> 5] |= 1u << (q & 31);
        else if (d[p] > 0.0)
          return fail(c, VPD_ERR_UNSUPPORTED, "Partial bond damage is not representable.");
      }
    HIPTRY(c, hipMemcpyAsync(c->broken, bits.data(), sizeof(uint32_t) * bits.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return VPD_OK;
  }
  if (field == VPD_FIELD_STRESS_N) {
    // Stored as the Voigt six the VUMAT sees (TensorVoigtMap.cpp:9-16 symmetrises).
    const double *t = static_cast<const double *>(host);
    std::vector<double> v6((size_t)c->n * 6);
    for (int64_t i = 0; i < c->n; ++i) {
      const double *f = t + 9 * i;
      const double vv[6] = {f[0], f[4], f[8], 0.5 * (f[5] + f[7]), 0.5 * (f[2] + f[6]),
                            0.5 * (f[1] + f[3])};
      for (int k = 0; k < 6; ++k)
        v6[(size_t)i * 6 + k] = vv[k];
    }
    return put_planes(c, c->sig, v6.data(), 6);
  }
  if (field == VPD_FIELD_STRESS || field == VPD_FIELD_STRAIN || field == VPD_FIELD_FORCE_DENSITY ||
      field == VPD_FIELD_DAMAGE_INDEX || field == VPD_FIELD_RETURN_MAP_ITERATIONS ||
      field == VPD_FIELD_REFERENCE_POSITION || field == VPD_FIELD_VOLUME)
    return fail(c, VPD_ERR_ARGUMENT, "Field is not writable.");
  double *dst = field_ptr(c, field);
  if (!dst)
    return fail(c, VPD_ERR_ARGUMENT, "Unknown field.");
  if (field == VPD_FIELD_DEFGRAD)
    c->haveFnew = true;
  if (field == VPD_FIELD_DEFGRAD_N && !c->props.empty())
    c->haveMaterial = true; // the caller supplies the material history
  if ((field == VPD_FIELD_VELOCITY || field == VPD_FIELD_ACCELERATION) && c->density > 0.0)
    c->haveDynamics = true;
  if (field == VPD_FIELD_SHAPE_INVERSE) {
    if (!c->haveNeighbors)
      return fail(c, VPD_ERR_STATE, "Neighbour lists have not been built.");
    c->haveShape = true;
  }
  if (field == VPD_FIELD_DISPLACEMENT || field == VPD_FIELD_VELOCITY ||
      field == VPD_FIELD_ACCELERATION)
    c->predicted = false;
  return put_planes(c, dst, static_cast<const double *>(host), components(c, field));
}

int vpd_download(vpd_context *c, int field, void *host) {
  if (int s = require_block(c))
    return s;
  if (!host)
    return fail(c, VPD_ERR_ARGUMENT, "Null host buffer.");
  if (int s = check_device(c))
    return s;
  if (field == VPD_FIELD_BOND_DAMAGE) {
    std::vector<uint32_t> bits(std::max<int64_t>(1, c->brokenWords));
    HIPTRY(c, hipMemcpyAsync(bits.data(), c->broken, sizeof(uint32_t) * bits.size(),
                             hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    double *d = static_cast<double *>(host);
    for (int64_t i = 0; i < c->n; ++i)
      for (int64_t p = c->hostRowptr[i]; p < c->hostRowptr[i + 1]; ++p) {
        const int64_t q = sell_slot(c, i, p);
        d[p] = (bits[q >> 5] >> (q & 31)) & 1u ? 1.0 : 0.0;
      }
    return VPD_OK;
  }
  if (field == VPD_FIELD_STRESS_N) {
    std::vector<double> v6((size_t)c->n * 6);
    if (int s = get_planes(c, c->sig, v6.data(), 6))
      return s;
    double *t = static_cast<double *>(host);
    for (int64_t i = 0; i < c->n; ++i) {
      const double *v = &v6[(size_t)i * 6];
      const double f[9] = {v[0], v[5], v[4], v[5], v[1], v[3], v[4], v[3], v[2]};
      std::memcpy(t + 9 * i, f, sizeof(f));
    }
    return VPD_OK;
  }
  if (field == VPD_FIELD_RETURN_MAP_ITERATIONS) {
    if (!c->iters || !c->haveFields)
      return fail(c, VPD_ERR_STATE, "Field is not available: run a step with want_fields.");
    HIPTRY(c, hipMemcpyAsync(host, c->iters, sizeof(int) * c->n, hipMemcpyDeviceToHost,
                             c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return VPD_OK;
  }
  if (field == VPD_FIELD_DAMAGE_INDEX) {
    if (!c->haveNeighbors)
      return fail(c, VPD_ERR_STATE, "Neighbour lists have not been built.");
    if (!c->dmgIndex && dalloc(c, c->dmgIndex, c->n))
      return VPD_ERR_DEVICE;
    k_damage_index<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, c->vol, sell_of(c), c->broken,
                                                         c->dmgIndex);
    HIPTRY(c, hipGetLastError());
  }
  if ((field == VPD_FIELD_STRESS || field == VPD_FIELD_STRAIN) && !c->haveFields)
    return fail(c, VPD_ERR_STATE, "Field is not available: run a step with want_fields.");
  if (field == VPD_FIELD_FORCE_DENSITY && !c->haveForce)
    return fail(c, VPD_ERR_STATE, "Field is not available: run a step with want_fields.");
  double *src = field_ptr(c, field);
  if (!src)
    return fail(c, VPD_ERR_ARGUMENT, "Unknown field.");
  return get_planes(c, src, static_cast<double *>(host), components(c, field));
}

int vpd_predict(vpd_context *c, double dt, const double ramp[3]) {
  if (int s = require_block(c))
    return s;
  if (!c->haveDynamics)
    return fail(c, VPD_ERR_STATE, "Invalid explicit dynamics predictor input.");
  if (!(dt > 0.0) || !ramp)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid explicit dynamics predictor input.");
  c->predicted = false;
  return launch_predict(c, dt, ramp);
}

int vpd_apply_constraints(vpd_context *c, const double ramp[3]) {
  if (int s = require_block(c))
    return s;
  if (!ramp)
    return fail(c, VPD_ERR_ARGUMENT, "Null ramp.");
  return timed(c, K_CONSTRAIN, [&] {
    k_constrain<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, planes(c->u, c->n),
                                                      planes(c->v, c->n), planes(c->a, c->n),
                                                      c->codes, ramp_of(ramp));
  });
}

int vpd_compute_deformation_gradient(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (!c->haveShape)
    return fail(c, VPD_ERR_STATE, "PD correspondence state is not initialized.");
  if (int s = launch_defgrad(c))
    return s;
  c->haveFnew = true;
  return check_device(c);
}

int vpd_update_material(vpd_context *c, double dt, double stepTime, double totalTime) {
  (void)stepTime;
  (void)totalTime; // passed through to the VUMAT, which the Johnson-Cook routine ignores
  if (int s = require_block(c))
    return s;
  if (!(dt > 0.0))
    return fail(c, VPD_ERR_ARGUMENT, "VUMAT-PD material update requires dt > 0.");
  if (!c->haveMaterial)
    return fail(c, VPD_ERR_STATE,
                "VUMAT material state is not initialized for this point block.");
  if (!c->haveFnew)
    return fail(c, VPD_ERR_STATE, "No deformation gradient to update from.");
  if (int s = ensure_outputs(c))
    return s;
  if (int s = launch_material(c, dt, true))
    return s;
  c->haveFields = true;
  return check_device(c, ~ERR_PIOLA_DET);
}

int vpd_compute_force_and_correct(vpd_context *c, double dt, const double ramp[3]) {
  if (int s = require_ready(c))
    return s;
  if (!(dt > 0.0) || !ramp)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid explicit dynamics corrector input.");
  if (int s = ensure_outputs(c))
    return s;
  if (int s = launch_force(c, dt, ramp, false, nullptr, true))
    return s;
  c->haveForce = true;
  return check_device(c);
}

int vpd_step(vpd_context *c, const vpd_step_args *g) {
  if (int s = require_ready(c))
    return s;
  if (!g || !(g->dt > 0.0))
    return fail(c, VPD_ERR_ARGUMENT, "Invalid explicit dynamics predictor input.");
  if (g->want_fields)
    if (int s = ensure_outputs(c))
      return s;
  if (!c->predicted)
    if (int s = launch_predict(c, g->dt, g->ramp))
      return s;
  if (int s = launch_defgrad(c))
    return s;
  if (int s = launch_material(c, g->dt, g->want_fields != 0))
    return s;
  if (int s = launch_force(c, g->dt, g->ramp, g->fuse_next != 0, g->ramp_next,
                           g->want_fields != 0))
    return s;
  c->predicted = g->fuse_next != 0;
  c->haveFnew = true;
  c->haveFields = g->want_fields != 0;
  c->haveForce = g->want_fields != 0;
  return VPD_OK;
}

int vpd_compute_force_density(vpd_context *c, const double *cauchy, double *force) {
  if (int s = require_block(c))
    return s;
  if (!cauchy || !force)
    return fail(c, VPD_ERR_ARGUMENT, "PD force-state inputs do not match particle count.");
  if (!c->haveShape)
    return fail(c, VPD_ERR_STATE, "PD correspondence state is not initialized.");
  if (!c->haveFnew)
    return fail(c, VPD_ERR_STATE, "No deformation gradient for the force state.");
  if (int s = ensure_outputs(c))
    return s;
  if (int s = put_planes(c, c->sigOut, cauchy, 9))
    return s;
  k_piola<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, planes(c->sigOut, c->n),
                                                planes(c->fnew, c->n), planes(c->kinv, c->n),
                                                planes(c->amap, c->n), c->derr);
  HIPTRY(c, hipGetLastError());
  if (int s = check_device(c))
    return s;
  ForceArgs g{};
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.x = planes(c->x, c->n);
  g.amap = planes(c->amap, c->n);
  g.forceOut = planes(c->force, c->n);
  g.vol = c->vol;
  g.bonds = sell_of(c);
  g.broken = c->broken;
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  if (int s = timed(c, K_FORCE, [&] {
        if (dmg)
          k_force<true, false, false><<<grid8_of(c->n), 256, 0, c->stream>>>(g);
        else
          k_force<false, false, false><<<grid8_of(c->n), 256, 0, c->stream>>>(g);
      }))
    return s;
  c->haveForce = true;
  return get_planes(c, c->force, force, 3);
}

int vpd_corotational_kinematics(const double *fN, const double *fNP1, const double *vN,
                                const double *rN, double dt, int64_t n, double *L, double *dbar,
                                double *relSpin, double *r1, double *v1) {
  if (!(dt > 0.0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "Corotational update requires dt > 0.");
  if (n <= 0)
    return VPD_OK;
  if (!fN || !fNP1 || !vN || !rN || !L || !dbar || !relSpin || !r1 || !v1)
    return fail(nullptr, VPD_ERR_ARGUMENT, "Null corotational argument.");
  vpd_context tmp;
  tmp.n = n;
  vpd_context *c = &tmp;
  double *buf = nullptr;
  int *err = nullptr;
  auto release = [&] {
    dfree(buf);
    dfree(err);
    if (tmp.stream)
      (void)hipStreamDestroy(tmp.stream);
    tmp.stream = nullptr;
  };
  if (hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, "HIP stream creation failed.");
  if (dalloc(c, buf, 9 * 9 * n) || dalloc(c, err, 1)) {
    release();
    return fail(nullptr, VPD_ERR_DEVICE, tmp.err);
  }
  double *p[9];
  for (int k = 0; k < 9; ++k)
    p[k] = buf + (int64_t)k * 9 * n;
  const double *in[4] = {fN, fNP1, vN, rN};
  double *out[5] = {L, dbar, relSpin, r1, v1};
  int st = VPD_OK;
  for (int k = 0; k < 4 && !st; ++k)
    st = put_planes(c, p[k], in[k], 9);
  if (!st && hipMemsetAsync(err, 0, sizeof(int), tmp.stream) != hipSuccess)
    st = VPD_ERR_DEVICE;
  if (!st) {
    k_corotational_full<<<grid_of(n), 256, 0, tmp.stream>>>(
        n, planes(p[0], n), planes(p[1], n), planes(p[2], n), planes(p[3], n), dt,
        planes(p[4], n), planes(p[5], n), planes(p[6], n), planes(p[7], n), planes(p[8], n), err);
    int bits = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bits, err, sizeof(int), hipMemcpyDeviceToHost, tmp.stream) != hipSuccess ||
        hipStreamSynchronize(tmp.stream) != hipSuccess)
      st = VPD_ERR_DEVICE;
    else if (bits) {
      release();
      return fail(nullptr, VPD_ERR_NUMERIC, "Corotational update encountered a singular tensor.");
    }
  }
  for (int k = 0; k < 5 && !st; ++k)
    st = get_planes(c, p[4 + k], out[k], 9);
  release();
  return st ? fail(nullptr, st, tmp.err) : VPD_OK;
}

int vpd_rotate_and_strain(const double *stress, const double *rotation, const double *defgrad,
                          int64_t n, double *spatialStress, double *strain) {
  if (n <= 0)
    return VPD_OK;
  vpd_context tmp;
  tmp.n = n;
  vpd_context *c = &tmp;
  double *buf = nullptr;
  auto release = [&] {
    dfree(buf);
    if (tmp.stream)
      (void)hipStreamDestroy(tmp.stream);
    tmp.stream = nullptr;
  };
  if (hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, "HIP stream creation failed.");
  if (dalloc(c, buf, 5 * 9 * n)) {
    release();
    return fail(nullptr, VPD_ERR_DEVICE, tmp.err);
  }
  double *p[5];
  for (int k = 0; k < 5; ++k)
    p[k] = buf + (int64_t)k * 9 * n;
  const bool rot = stress && rotation && spatialStress, gl = defgrad && strain;
  int st = VPD_OK;
  if (rot)
    st = put_planes(c, p[0], stress, 9) || put_planes(c, p[1], rotation, 9);
  if (!st && gl)
    st = put_planes(c, p[2], defgrad, 9);
  if (!st) {
    k_rotate_gl<<<grid_of(n), 256, 0, tmp.stream>>>(
        n, planes(p[0], n), planes(p[1], n), planes(rot ? p[3] : nullptr, n), planes(p[2], n),
        planes(gl ? p[4] : nullptr, n));
    if (hipGetLastError() != hipSuccess)
      st = VPD_ERR_DEVICE;
  }
  if (!st && rot)
    st = get_planes(c, p[3], spatialStress, 9);
  if (!st && gl)
    st = get_planes(c, p[4], strain, 9);
  release();
  return st ? fail(nullptr, st, tmp.err) : VPD_OK;
}

int vpd_correct_velocity(vpd_context *c, double dt, const double *force) {
  if (int s = require_block(c))
    return s;
  if (!(dt > 0.0) || !force)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid explicit dynamics corrector input.");
  if (!c->haveDynamics)
    return fail(c, VPD_ERR_STATE, "Invalid explicit dynamics corrector input.");
  if (int s = ensure_outputs(c))
    return s;
  if (int s = put_planes(c, c->force, force, 3))
    return s;
  k_correct<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, planes(c->v, c->n), planes(c->a, c->n),
                                                  planes(c->force, c->n), dt, c->density);
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int vpd_update_damage(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (!c->haveNeighbors)
    return fail(c, VPD_ERR_STATE, "Neighbour lists have not been built.");
  if (c->damageMode != VPD_DAMAGE_CRITICAL_STRETCH)
    return VPD_OK;
  const int useCompression = c->failure > 0.0 && std::isfinite(c->failure);
  k_bond_damage<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, planes(c->x, c->n),
                                                      planes(c->u, c->n), sell_of(c), c->broken,
                                                      c->initiation, c->failure, useCompression);
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int vpd_synchronize(vpd_context *c) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return check_device(c);
}

int vpd_set_timing(vpd_context *c, int enable) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  c->timing = enable != 0;
  return VPD_OK;
}

int vpd_collect_timing(vpd_context *c, vpd_kernel_time *out, int capacity, int *count) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  HIPTRY(c, hipStreamSynchronize(c->stream));
  double total[K_COUNT] = {0};
  int64_t launches[K_COUNT] = {0};
  for (auto &r : c->recs) {
    float ms = 0.f;
    HIPTRY(c, hipEventElapsedTime(&ms, r.start, r.stop));
    total[r.kernel] += ms;
    launches[r.kernel] += 1;
    c->eventPool.push_back(r.start);
    c->eventPool.push_back(r.stop);
  }
  c->recs.clear();
  int k = 0;
  for (int id = 0; id < K_COUNT; ++id)
    if (launches[id] && k < capacity && out) {
      out[k].name = kKernelNames[id];
      out[k].total_ms = total[id];
      out[k].launches = launches[id];
      ++k;
    }
  if (count)
    *count = k;
  return VPD_OK;
}

int vpd_device_malloc(uint64_t bytes, void **out) {
  if (!out)
    return fail(nullptr, VPD_ERR_ARGUMENT, "Null output pointer.");
  *out = nullptr;
  hipError_t e = hipMalloc(out, bytes ? bytes : 1);
  return e == hipSuccess ? VPD_OK
                         : fail(nullptr, VPD_ERR_DEVICE,
                                std::string("HIP error: ") + hipGetErrorString(e));
}

int vpd_device_free(void *ptr) {
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? VPD_OK
                         : fail(nullptr, VPD_ERR_DEVICE,
                                std::string("HIP error: ") + hipGetErrorString(e));
}

int vpd_memcpy(void *dst, const void *src, uint64_t bytes, int to_device) {
  hipError_t e = hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice
                                                      : hipMemcpyDeviceToHost);
  return e == hipSuccess ? VPD_OK
                         : fail(nullptr, VPD_ERR_DEVICE,
                                std::string("HIP error: ") + hipGetErrorString(e));
}

int vpd_vumat_adapter_call_device(int *nblock, int *ndir, int *nshr, int *nstatev, int *nfieldv,
                                  int *nprops, int *lanneal, double *stepTime, double *totalTime,
                                  double *dt, char *cmname, double *coordMp, double *charLength,
                                  double *props, double *density, double *strainInc,
                                  double *relSpinInc, double *tempOld, double *stretchOld,
                                  double *defgradOld, double *fieldOld, double *stressOld,
                                  double *stateOld, double *enerInternOld, double *enerInelasOld,
                                  double *tempNew, double *stretchNew, double *defgradNew,
                                  double *fieldNew, double *stressNew, double *stateNew,
                                  double *enerInternNew, double *enerInelasNew) {
  (void)lanneal;
  (void)stepTime;
  (void)totalTime;
  (void)cmname;
  (void)coordMp;
  (void)charLength;
  (void)density;
  (void)relSpinInc;
  if (!nblock || !ndir || !nshr || !nstatev || !nfieldv || !nprops || !dt || !props)
    return fail(nullptr, VPD_ERR_ARGUMENT, "Null VUMAT argument.");
  if (*ndir != 3 || *nshr != 3)
    return fail(nullptr, VPD_ERR_ARGUMENT, "JohnsonCook VUMAT requires ndir=3,nshr=3");
  if (*nprops < 14 || *nstatev < 6)
    return fail(nullptr, VPD_ERR_ARGUMENT, "JohnsonCook VUMAT requires 14 props and 6 states");
  if (*nblock <= 0)
    return VPD_OK;
  std::vector<double> hp(*nprops);
  if (hipMemcpy(hp.data(), props, sizeof(double) * hp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, "Reading VUMAT props from the device failed.");
  VumatArgs g;
  g.nblock = *nblock;
  g.nstatev = *nstatev;
  g.nfieldv = *nfieldv;
  g.jc = make_jc(hp);
  g.dt = *dt;
  g.strainInc = strainInc;
  g.tempOld = tempOld;
  g.stretchOld = stretchOld;
  g.defgradOld = defgradOld;
  g.fieldOld = fieldOld;
  g.stressOld = stressOld;
  g.stateOld = stateOld;
  g.enerInternOld = enerInternOld;
  g.enerInelasOld = enerInelasOld;
  g.tempNew = tempNew;
  g.stretchNew = stretchNew;
  g.defgradNew = defgradNew;
  g.fieldNew = fieldNew;
  g.stressNew = stressNew;
  g.stateNew = stateNew;
  g.enerInternNew = enerInternNew;
  g.enerInelasNew = enerInelasNew;
  k_vumat_jc<<<grid_of(g.nblock), 256>>>(g);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
  return VPD_OK;
}

} // extern "C"
