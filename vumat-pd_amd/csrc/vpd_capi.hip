// C ABI of the MI355X PD increment (include/vpd.h): context lifetime, device buffers,
// kernel launches on the context stream, state transfer and error mapping.
#include "vpd.h"
#include "vpd_kernels.hpp"
#include "vpd_mlp.hpp"
#include "vpd_neighbors.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
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
  K_MLP_SEED,
  K_RETURN_MAP,
  K_COUNT
};
const char *kKernelNames[K_COUNT] = {"predict",      "constrain", "defgrad_bond_pass",
                                     "material_jc", "force_bond_pass", "shape_tensor",
                                     "neighbor_build", "mlp_seed_mfma", "return_map"};

struct TimingRec {
  int kernel;
  hipEvent_t start, stop;
};

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

} // namespace

struct vpd_mlp {
  hipStream_t stream = nullptr;
  int maxB = 0;
  float *theta = nullptr, *norm = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  float *xb = nullptr, *yb = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *y = nullptr;
  float *dy = nullptr, *dz3 = nullptr, *dz2 = nullptr, *dz1 = nullptr, *part = nullptr;
  float *sq = nullptr, *xs = nullptr, *ys = nullptr;
  int *order = nullptr;
  int64_t nData = 0, cursor = 0;
  int adamStep = 0;
  // bf16 MFMA training (VPD_MLP_TRAIN_MFMA): feature-major layer inputs and gradients over
  // the padded batch, dy, and the per-chunk gradient partials.
  int form = VPD_MLP_TRAIN_MFMA;
  int64_t ld = 0;
  short *xT = nullptr, *h1T = nullptr, *h2T = nullptr, *h3T = nullptr, *dz1T = nullptr,
        *dz2T = nullptr, *dz3T = nullptr;
  float *dyT = nullptr, *gpart = nullptr;
  double *seedBuf = nullptr; // vpd_return_map_learned_device: the seed per point (on demand)
  int64_t seedCap = 0;
  std::string err;
};

struct vpd_context {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  // configuration
  double horizon = 0, density = 0, charLength = 1, tref = 293, materialDensity = 0;
  std::vector<double> props, initialState;
  int nstatev = 0, damageMode = 0, returnMap = 0;
  double initiation = 0, failure = 1;
  int materialModel = VPD_MATERIAL_JOHNSON_COOK;
  JcParams jc{};
  ElasticParams elastic{};
  // block
  int64_t n = 0, nnz = 0;
  double *x = nullptr, *vol = nullptr;
  int *rowptr = nullptr, *col = nullptr;
  uint8_t *codes = nullptr;
  double *kinv = nullptr;
  double *u = nullptr, *v = nullptr, *a = nullptr;
  double *fN = nullptr, *vN = nullptr, *rN = nullptr, *sig = nullptr, *statev = nullptr;
  double *temp = nullptr, *eint = nullptr, *einel = nullptr;
  double *fnew = nullptr, *amap = nullptr; // amap: A = P K⁻¹ as 5 component pairs (10 n)
  double *xv = nullptr;                     // (X_x, X_y), (X_z, V) pairs (4 n), static
  double *sigOut = nullptr, *strainOut = nullptr, *force = nullptr, *dmgIndex = nullptr;
  int *iters = nullptr;
  uint32_t *broken = nullptr; // one bit per sliced bond slot
  // One byte per particle, 1 when any of its bonds may be broken: set by every kernel that
  // breaks a bond, rebuilt when the flags are uploaded, cleared with them.
  uint8_t *anyBroken = nullptr;
  int64_t brokenWords = 0;
  int *colS = nullptr;         // sliced bond layout (see Sell in vpd_kernels.hpp)
  int64_t *sliceBase = nullptr;
  int64_t sellTotal = 0;
  std::vector<int> hostRowptr;
  std::vector<int64_t> hostBase;
  int *derr = nullptr;
  bool haveNeighbors = false, haveShape = false, haveMaterial = false, haveDynamics = false;
  bool haveFnew = false, haveFields = false, haveForce = false, predicted = false;
  // F_{n+1} of an eager single-specimen step is not copied into fN: the material stage leaves
  // it in fnew and the two buffers swap (swapF asks for that on the next material stage;
  // fInN: the latest F, which VPD_FIELD_DEFGRAD returns, is in fN since the swap).
  bool swapF = false, fInN = false;
  bool timing = false;
  // With timing on, the kernels of every timingStride-th step are timed (timingTick counts
  // steps; timeThis says whether the step being taken is one of them).
  int timingStride = 1;
  int64_t timingTick = 0;
  bool timeThis = true;
  // LDS bytes reserved per 256-thread workgroup of each bond pass: a cap on how many
  // workgroups share a CU, and with it on the span of particles whose neighbour data an
  // XCD's L2 must hold at once. Chosen by vpd_autotune.
  int ldsDefgrad = 0, ldsForce = 0;
  // Sweeps per bond pass (see Rounds in vpd_kernels.hpp). Chosen by vpd_autotune.
  int roundsDefgrad = 1, roundsForce = 1;
  // Window of the windowed gathers per bond pass (particles staged per window; 0: direct
  // gathers) and the most bonds a lane takes from one window (0: no limit). Chosen by
  // vpd_autotune.
  int winDefgrad = 0, winForce = 0;
  int takeDefgrad = 0, takeForce = 0;
  // Form of the windowed gathers per bond pass: 0 walks the bond lists, 1 replays a window
  // plan, 2 replays it with the next window prefetched (WindowPlan in vpd_kernels.hpp), 3
  // replays a plan of windows shared by each workgroup (GroupPlan), 4 shares each particle's
  // row among a group of lanes (k_defgrad_m, k_force_m). Chosen by vpd_autotune.
  int modeDefgrad = 0, modeForce = 0;
  // The settings of each bond pass that the last vpd_autotune timed again in its final stage,
  // fastest first (vpd_autotune_candidates).
  struct TunedSetting {
    int lds, rounds, window, take, mode;
    float ms;
  };
  std::vector<TunedSetting> tunedTop[2];
  // The window plan of each bond pass, built for the pass's window and bond limit when first
  // needed after the bond lists or those knobs change.
  struct PlanBuffers {
    int *first = nullptr, *start = nullptr;
    uint64_t *rec = nullptr;
    int64_t capacity = 0; // windows `start` and `rec` hold
    int window = 0, take = 0;
    bool valid = false;
  } plan[2];
  // The group-window plan of each bond pass, kept like `plan`.
  struct GroupBuffers {
    int *first = nullptr, *meta = nullptr;
    uint64_t *off = nullptr;
    int64_t windows = 0, blocks = 0; // what `meta` (windows) and `off` (entry blocks) hold
    int window = 0;
    bool valid = false;
  } gplan[2];
  // Return map of the material stage; the learned mode adds the network, the strain
  // increments it reads and the seeds it writes.
  double rtol = 1.0e-12;
  vpd_mlp *mlp = nullptr;
  double *inc = nullptr, *seed = nullptr;
  int *fallbacks = nullptr;
  // The compacted material update: list lengths, particle per slot, the network's slots,
  // per slot the strain increments and the first iteration's 1 − θᵐ and yield. listZeroed: the lengths are known to be zero (a force pass
  // cleared them after the last material update), so the next update need not clear them.
  int *listCount = nullptr, *list = nullptr;
  double *incList = nullptr;
  bool listZeroed = false, capturing = false;
  std::vector<TimingRec> recs;
  std::vector<hipEvent_t> eventPool;
  // Run diagnostics (vpd_enable_diagnostics): per-chunk partial rows, the step records kept
  // until fetched, the observation weights and lateral face flags, and the strain-increment
  // scratch of the split (learned) material update.
  bool diagOn = false;
  DiagArgs diag{};
  double *diagPartials = nullptr, *diagRecords = nullptr, *diagW = nullptr;
  double *diagDstrain = nullptr;
  uint8_t *diagLat = nullptr;
  int diagCapacity = 0, diagCount = 0;
  // Batched specimens: the chunks the diagnostics reduce (each within one specimen, laid out
  // from its first particle) and each specimen's first chunk; records per step = specimens.
  int64_t *diagChunkFirst = nullptr, *diagChunkEnd = nullptr, *diagSpecChunk = nullptr;
  int64_t diagChunks = 0;
  int diagPerStep = 1;
  // Batched specimens (vpd_set_specimens): specimen of each particle, each specimen's time
  // step, end velocity and last step, and the step being taken. batch.id stays null for a
  // block that is one specimen.
  Batch batch{};
  uint16_t *specId = nullptr;
  double *specDt = nullptr, *specVel = nullptr;
  int *specLast = nullptr;
  // The two step-counter slots of graph-replayed steps (a captured pair of steps reads one
  // each), and the slot the force pass being captured writes the next step's number to.
  int *stepDev = nullptr, *stepNextDev = nullptr;
  double *specBand = nullptr; // vpd_set_specimen_bands: prescribed band rows
  int64_t *specBandOff = nullptr;
  int specCount = 0;
  std::vector<int64_t> specFirst;
  std::vector<int32_t> specLastHost;
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
  const bool on = c->timing && c->timeThis;
  if (on) {
    r.start = take_event(c);
    r.stop = take_event(c);
    HIPTRY(c, hipEventRecord(r.start, c->stream));
  }
  launch();
  HIPTRY(c, hipGetLastError());
  if (on) {
    HIPTRY(c, hipEventRecord(r.stop, c->stream));
    c->recs.push_back(r);
  }
  return VPD_OK;
}

// For the duration of one step: whether its kernels are timed (every timingStride-th step).
struct StepTiming {
  vpd_context *c;
  explicit StepTiming(vpd_context *ctx) : c(ctx) {
    c->timeThis = c->timingTick++ % c->timingStride == 0;
  }
  ~StepTiming() { c->timeThis = true; }
};

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
PairPlanes pairs(double *p, int64_t n) { return PairPlanes{reinterpret_cast<d2v *>(p), n}; }

int grid_of(int64_t n) { return (int)ceil_div(n, 256); }
// Grids of the XCD-remapped kernels are rounded up to whole rounds of 8 workgroups.
int grid8_of(int64_t n) { return (int)(ceil_div(ceil_div(n, 256), 8) * 8); }

Rounds rounds_of(int64_t n, int rounds) {
  const int64_t blocks = ceil_div(n, 256);
  Rounds r;
  r.n = n;
  r.rounds = (int)std::max<int64_t>(1, std::min<int64_t>(rounds, blocks));
  r.blocksPerRound = ceil_div(blocks, r.rounds);
  return r;
}
int grid_rounds(const Rounds &r) { return (int)(ceil_div(r.blocksPerRound, 8) * 8); }

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
    return c->fInN ? c->fN : c->fnew;
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

// Host displacement rows <-> the device pair layout (u_x, u_y) planes, then (u_z, 0).
int put_u(vpd_context *c, const double *host) {
  const int64_t n = c->n;
  std::vector<double> stage((size_t)4 * n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    stage[2 * i] = host[3 * i];
    stage[2 * i + 1] = host[3 * i + 1];
    stage[2 * (n + i)] = host[3 * i + 2];
  }
  HIPTRY(c, hipMemcpyAsync(c->u, stage.data(), sizeof(double) * stage.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int get_u(vpd_context *c, double *host) {
  const int64_t n = c->n;
  std::vector<double> stage((size_t)4 * n);
  HIPTRY(c, hipMemcpyAsync(stage.data(), c->u, sizeof(double) * stage.size(),
                           hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) {
    host[3 * i] = stage[2 * i];
    host[3 * i + 1] = stage[2 * i + 1];
    host[3 * i + 2] = stage[2 * (n + i)];
  }
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
// The scan itself in caller-supplied work buffers: part (count int64) and bt
// (ceil(count / SCAN_ITEMS) + 1 int64). Synchronises the stream; *total receives the sum, and
// `out` is written only when the sum fits int32 (otherwise *total > INT_MAX tells the caller).
hipError_t scan_ints_in(hipStream_t st, const int *in, int *out, int64_t count, int64_t *part,
                        int64_t *bt, int64_t *total) {
  const int64_t nb = ceil_div(count, SCAN_ITEMS);
  k_scan_block<<<(int)nb, 256, 0, st>>>(count, in, part, bt);
  k_scan_totals<<<1, 256, 0, st>>>(nb, bt, bt + nb);
  hipError_t e = hipMemcpyAsync(total, bt + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipStreamSynchronize(st);
  if (e != hipSuccess || *total > INT_MAX)
    return e;
  k_scan_add<<<grid_of(count), 256, 0, st>>>(count, part, bt, out);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipStreamSynchronize(st);
  return e;
}

int scan_ints(vpd_context *c, const int *in, int *out, int64_t count, int64_t *total) {
  const int64_t nb = ceil_div(count, SCAN_ITEMS);
  int64_t *part = nullptr, *bt = nullptr;
  if (dalloc(c, part, count) || dalloc(c, bt, nb + 1))
    return VPD_ERR_DEVICE;
  const hipError_t e = scan_ints_in(c->stream, in, out, count, part, bt, total);
  dfree(part);
  dfree(bt);
  HIPTRY(c, e);
  if (*total > INT_MAX)
    return fail(c, VPD_ERR_UNSUPPORTED, "Neighbour count exceeds the int32 CSR range.");
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
  if (!c->fallbacks && dalloc(c, c->fallbacks, c->n))
    return VPD_ERR_DEVICE;
  return VPD_OK;
}

// Sliced copy of the bond lists: slice s holds particles 64s..64s+63, padded to its
// longest row; bond flags are indexed by sliced slot.
int build_sell(vpd_context *c) {
  const int64_t n = c->n;
  for (auto &p : c->plan) { // plans follow the bond lists (and the particle count)
    dfree(p.first);
    p.valid = false;
  }
  for (auto &p : c->gplan) {
    dfree(p.first);
    p.valid = false;
  }
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
    c->hostBase[s + 1] = c->hostBase[s] + 64 * (int64_t)((m + 3) / 4 * 4);
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
  if (dalloc(c, c->anyBroken, n))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemsetAsync(c->anyBroken, 0, std::max<int64_t>(1, n), c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

// Sliced slot of CSR entry p of row i.
int64_t sell_slot(const vpd_context *c, int64_t i, int64_t p) {
  return vpd::sell_slot(c->hostBase[i >> 6], i, p - c->hostRowptr[i]);
}

Sell sell_of(const vpd_context *c) { return Sell{c->sliceBase, c->colS, c->rowptr}; }

UPairs upairs(const vpd_context *c) { return UPairs{c->u, c->n}; }

Ramp ramp_of(const double r[3]) { return Ramp{r[0], r[1], r[2]}; }

int launch_predict(vpd_context *c, double dt, const double ramp[3]) {
  return timed(c, K_PREDICT, [&] {
    k_predict<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, upairs(c), planes(c->v, c->n),
                                                    planes(c->a, c->n), c->codes, dt,
                                                    ramp_of(ramp), c->batch);
  });
}

// LDS of a windowed bond pass: 4 wavefronts × `pairs` component pairs × W particles. The
// window sizes are compile-time (the LDS offsets of the component planes become immediates).
constexpr int kMaxLds = 163840;
constexpr int kWindows[] = {64, 80, 96, 128, 160};
int window_lds(int pairs, int window) { return 4 * pairs * window * 16; }
bool window_supported(int window) {
  for (int w : kWindows)
    if (w == window)
      return true;
  return false;
}

// f.template operator()<W>() for the compile-time window equal to `window`.
template <class F> void with_window(int window, F &&f) {
  switch (window) {
  case 64:
    f(std::integral_constant<int, 64>{});
    break;
  case 80:
    f(std::integral_constant<int, 80>{});
    break;
  case 96:
    f(std::integral_constant<int, 96>{});
    break;
  case 128:
    f(std::integral_constant<int, 128>{});
    break;
  default:
    f(std::integral_constant<int, 160>{});
    break;
  }
}
int launch_lds(int cap, int pairs, int window) {
  return window > 0 ? std::max(cap, window_lds(pairs, window)) : cap;
}
// Dynamic LDS of a prefetching planned pass: its two window buffers are static, so only the
// rest of the reservation (an occupancy cap above them) is asked for.
int prefetch_lds(int cap, int pairs, int window) {
  return std::max(0, cap - 2 * window_lds(pairs, window));
}
// The kernels' per-window bond limit for a tuning value (0: none).
int take_arg(int take) { return take > 0 ? take : INT_MAX; }
constexpr int kMaxTake = 1024;

// Planned windows (WindowPlan): window sizes with planned kernels, and a compile-time window.
constexpr int kPlannedWindows[] = {64, 80, 96, 128};
bool planned_supported(int window, int take) {
  if (take < 1 || take > kPlanTake)
    return false;
  for (int w : kPlannedWindows)
    if (w == window)
      return true;
  return false;
}
template <class F> void with_planned_window(int window, F &&f) {
  switch (window) {
  case 64:
    f(std::integral_constant<int, 64>{});
    break;
  case 80:
    f(std::integral_constant<int, 80>{});
    break;
  case 96:
    f(std::integral_constant<int, 96>{});
    break;
  default:
    f(std::integral_constant<int, 128>{});
    break;
  }
}

// Group windows (GroupPlan): window sizes with group-window kernels, and a compile-time
// window. The workgroup holds one window: `pairs` planes of W pairs (the force pass's 7
// planes of 640 pairs take 70 KB, two workgroups per CU), within the CU's 160 KB.
constexpr int kGroupWindows[] = {256, 384, 480, 512, 640};
// Blocks up to this many particles also try shared rows (gather form 4) in the autotuner,
// with each of these lanes per particle.
constexpr int64_t kRowBlockMax = 262144;
constexpr int kRowLanes[] = {16, 32, 64};
bool row_lanes_supported(int lanes) { return lanes == 16 || lanes == 32 || lanes == 64; }
template <class F> void with_row_lanes(int lanes, F &&f) {
  if (lanes == 16)
    f(std::integral_constant<int, 16>{});
  else if (lanes == 32)
    f(std::integral_constant<int, 32>{});
  else
    f(std::integral_constant<int, 64>{});
}
constexpr int kMaxGroupLds = kMaxLds;
bool group_window_supported(int window) {
  for (int w : kGroupWindows)
    if (w == window)
      return true;
  return false;
}
int group_lds(int pairs, int window) { return pairs * window * 16; }
template <class F> void with_group_window(int window, F &&f) {
  switch (window) {
  case 256:
    f(std::integral_constant<int, 256>{});
    break;
  case 384:
    f(std::integral_constant<int, 384>{});
    break;
  case 480:
    f(std::integral_constant<int, 480>{});
    break;
  case 640:
    f(std::integral_constant<int, 640>{});
    break;
  default:
    f(std::integral_constant<int, 512>{});
    break;
  }
}

// How a bond pass gathers its neighbour data: window 0 (direct) or W, the bond limit per
// window, the form (0 cursor walk, 1 planned, 2 planned with prefetch, 3 group windows, 4
// shared rows, whose "window" is the lanes per particle) and the plan of the form.
struct Gather {
  int window = 0, take = 0, mode = 0;
  WindowPlan plan{};
  GroupPlan gplan{};
};

// Builds the window plan of bond pass `pass` (0 F, 1 force) for (window, take), unless it is
// already there: count the windows of every wavefront, lay them out, fill them.
int build_plan(vpd_context *c, int pass, int window, int take) {
  auto &p = c->plan[pass];
  if (p.valid && p.first && p.window == window && p.take == take)
    return VPD_OK;
  if (!planned_supported(window, take))
    return fail(c, VPD_ERR_ARGUMENT, "Planned windows need a window of 64, 80, 96 or 128 "
                                     "particles and 1 to 8 bonds per window.");
  p.valid = false;
  const int64_t n = c->n, waves = ceil_div(n, 64);
  int *count = nullptr;
  if (dalloc(c, count, waves) || (!p.first && dalloc(c, p.first, waves + 1)))
    return VPD_ERR_DEVICE;
  with_planned_window(window, [&](auto w) {
    k_window_plan<decltype(w)::value, false><<<grid_of(n), 256, 0, c->stream>>>(
        n, sell_of(c), take, count, nullptr, nullptr, nullptr);
  });
  HIPTRY(c, hipGetLastError());
  std::vector<int> hostCount(waves), first(waves + 1, 0);
  HIPTRY(c, hipMemcpyAsync(hostCount.data(), count, sizeof(int) * waves, hipMemcpyDeviceToHost,
                           c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  dfree(count);
  int64_t total = 0;
  for (int64_t w = 0; w < waves; ++w) {
    first[w] = (int)total;
    total += hostCount[w];
    if (total >= INT_MAX)
      return fail(c, VPD_ERR_ARGUMENT, "Too many windows for a window plan.");
  }
  first[waves] = (int)total;
  if (total > p.capacity) {
    if (dalloc(c, p.start, total) || dalloc(c, p.rec, total * 64))
      return VPD_ERR_DEVICE;
    p.capacity = total;
  }
  HIPTRY(c, hipMemcpyAsync(p.first, first.data(), sizeof(int) * (waves + 1),
                           hipMemcpyHostToDevice, c->stream));
  with_planned_window(window, [&](auto w) {
    k_window_plan<decltype(w)::value, true><<<grid_of(n), 256, 0, c->stream>>>(
        n, sell_of(c), take, nullptr, p.first, p.start, p.rec);
  });
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  p.window = window;
  p.take = take;
  p.valid = true;
  return VPD_OK;
}

WindowPlan plan_of(const vpd_context *c, int pass) {
  const auto &p = c->plan[pass];
  return WindowPlan{p.first, p.start, p.rec};
}

// Builds the group-window plan of bond pass `pass` for `window`, unless it is already there:
// count the windows and entry blocks of every workgroup, lay them out, fill them.
int build_group_plan(vpd_context *c, int pass, int window) {
  auto &p = c->gplan[pass];
  if (p.valid && p.first && p.window == window)
    return VPD_OK;
  if (!group_window_supported(window))
    return fail(c, VPD_ERR_ARGUMENT,
                "Group windows need a window of 256, 384, 480, 512 or 640 particles and no "
                "bond limit per window (0).");
  p.valid = false;
  const int64_t n = c->n, groups = ceil_div(n, 256);
  int *count = nullptr, *firstBlock = nullptr;
  if (dalloc(c, count, 2 * groups) || (!p.first && dalloc(c, p.first, groups + 1)))
    return VPD_ERR_DEVICE;
  with_group_window(window, [&](auto w) {
    k_group_plan<decltype(w)::value, false><<<groups, 256, 0, c->stream>>>(
        n, sell_of(c), count, nullptr, nullptr, nullptr, nullptr);
  });
  HIPTRY(c, hipGetLastError());
  std::vector<int> hostCount(2 * groups), first(groups + 1, 0), firstBlk(groups, 0);
  HIPTRY(c, hipMemcpyAsync(hostCount.data(), count, sizeof(int) * 2 * groups,
                           hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  dfree(count);
  int64_t windows = 0, blocks = 0;
  for (int64_t g = 0; g < groups; ++g) {
    first[g] = (int)windows;
    firstBlk[g] = (int)blocks;
    windows += hostCount[2 * g];
    blocks += hostCount[2 * g + 1];
    if (windows >= INT_MAX / kGroupMeta || blocks >= INT_MAX)
      return fail(c, VPD_ERR_ARGUMENT, "Too many windows for a group-window plan.");
  }
  first[groups] = (int)windows;
  if (windows > p.windows) {
    if (dalloc(c, p.meta, kGroupMeta * windows))
      return VPD_ERR_DEVICE;
    p.windows = windows;
  }
  if (blocks > p.blocks) {
    if (dalloc(c, p.off, 256 * blocks))
      return VPD_ERR_DEVICE;
    p.blocks = blocks;
  }
  if (dalloc(c, firstBlock, groups))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemcpyAsync(p.first, first.data(), sizeof(int) * (groups + 1),
                           hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipMemcpyAsync(firstBlock, firstBlk.data(), sizeof(int) * groups,
                           hipMemcpyHostToDevice, c->stream));
  with_group_window(window, [&](auto w) {
    k_group_plan<decltype(w)::value, true><<<groups, 256, 0, c->stream>>>(
        n, sell_of(c), nullptr, p.first, firstBlock, p.meta, p.off);
  });
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  dfree(firstBlock);
  p.window = window;
  p.valid = true;
  return VPD_OK;
}

GroupPlan group_plan_of(const vpd_context *c, int pass) {
  const auto &p = c->gplan[pass];
  return GroupPlan{p.first, p.meta, p.off};
}

// Builds the plan a gather needs (none for direct and walked gathers; a walked gather is
// checked for a supported window).
int ensure_plan(vpd_context *c, int pass, const Gather &gx) {
  if (gx.mode == 4 || (gx.window > 0 && gx.window < 64)) {
    if (!row_lanes_supported(gx.window) || gx.mode != 4 || gx.take != 0)
      return fail(c, VPD_ERR_ARGUMENT, "Shared rows (form 4) go with a window of 16, 32 or 64 "
                                       "(lanes per particle) and no bond limit per window (0); "
                                       "windows 16 and 32 only with them.");
    return VPD_OK;
  }
  if (gx.window > 0 && gx.mode == 0 && !window_supported(gx.window))
    return fail(c, VPD_ERR_ARGUMENT, "Walked windows need a window of 64, 80, 96, 128 or 160 "
                                     "particles.");
  if (gx.mode == 3) {
    if (gx.take != 0)
      return fail(c, VPD_ERR_ARGUMENT,
                  "Group windows need a window of 256, 384, 480, 512 or 640 particles and no "
                  "bond limit per window (0).");
    if (group_lds(pass == 0 ? 4 : 7, gx.window) > kMaxGroupLds)
      return fail(c, VPD_ERR_ARGUMENT, "Group window exceeds the LDS of a workgroup.");
    return build_group_plan(c, pass, gx.window);
  }
  return gx.mode > 0 ? build_plan(c, pass, gx.window, gx.take) : VPD_OK;
}

// The plan pointers of a gather whose plan has been built.
Gather with_plan(const vpd_context *c, int pass, Gather gx) {
  if (gx.mode == 3)
    gx.gplan = group_plan_of(c, pass);
  else if (gx.mode > 0)
    gx.plan = plan_of(c, pass);
  return gx;
}

// The gather of bond pass `pass` as the knobs set it.
Gather gather_of(const vpd_context *c, int pass) {
  Gather gx;
  gx.window = pass == 0 ? c->winDefgrad : c->winForce;
  gx.take = pass == 0 ? c->takeDefgrad : c->takeForce;
  gx.mode = gx.window > 0 ? (pass == 0 ? c->modeDefgrad : c->modeForce) : 0;
  return with_plan(c, pass, gx);
}

// Builds the plans the knobs ask for (outside any graph capture: this allocates and waits).
int prepare_plans(vpd_context *c) {
  for (int pass = 0; pass < 2; ++pass)
    if (int s = ensure_plan(c, pass, gather_of(c, pass)))
      return s;
  return VPD_OK;
}

template <bool DAMAGE>
void launch_defgrad_t(vpd_context *c, int lds, const Rounds &rd, int useCompression,
                      const Batch &batch, const Gather &gx) {
  const int window = gx.window;
  if (window > 0 && gx.mode == 4) {
    with_row_lanes(window, [&](auto w) {
      constexpr int R = decltype(w)::value;
      k_defgrad_m<DAMAGE, R><<<(int)ceil_div(rd.n, 256 / R), 256, lds, c->stream>>>(
          rd.n, planes(c->x, c->n), pairs(c->xv, c->n), upairs(c), sell_of(c),
          planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->anyBroken,
          c->initiation, c->failure, useCompression, batch);
    });
    return;
  }
  if (window > 0 && gx.mode == 3) {
    with_group_window(window, [&](auto w) {
      k_defgrad_g<DAMAGE, decltype(w)::value>
          <<<grid_rounds(rd), 256, std::max(lds, group_lds(4, window)), c->stream>>>(
              rd, planes(c->x, c->n), pairs(c->xv, c->n), upairs(c), sell_of(c),
              planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->anyBroken,
              c->initiation, c->failure, useCompression, batch, gx.gplan);
    });
    return;
  }
  if (window > 0 && gx.mode > 0) {
    with_planned_window(window, [&](auto w) {
      constexpr int W = decltype(w)::value;
      if (gx.mode == 2)
        k_defgrad_p<DAMAGE, W, true><<<grid_rounds(rd), 256, prefetch_lds(lds, 4, window),
                                        c->stream>>>(
            rd, planes(c->x, c->n), pairs(c->xv, c->n), upairs(c), sell_of(c),
            planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->anyBroken,
            c->initiation, c->failure, useCompression, batch, gx.plan);
      else
        k_defgrad_p<DAMAGE, W, false><<<grid_rounds(rd), 256, launch_lds(lds, 4, window),
                                         c->stream>>>(
            rd, planes(c->x, c->n), pairs(c->xv, c->n), upairs(c), sell_of(c),
            planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->anyBroken,
            c->initiation, c->failure, useCompression, batch, gx.plan);
    });
    return;
  }
  if (window > 0) {
    with_window(window, [&](auto w) {
      k_defgrad_w<DAMAGE, decltype(w)::value>
          <<<grid_rounds(rd), 256, launch_lds(lds, 4, window), c->stream>>>(
              rd, planes(c->x, c->n), pairs(c->xv, c->n), upairs(c), sell_of(c),
              planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->anyBroken,
              c->initiation, c->failure, useCompression, batch, take_arg(gx.take));
    });
    return;
  }
  k_defgrad<DAMAGE><<<grid_rounds(rd), 256, lds, c->stream>>>(
      rd, planes(c->x, c->n), pairs(c->xv, c->n), upairs(c), sell_of(c),
      planes(c->kinv, c->n), planes(c->fnew, c->n), c->broken, c->anyBroken, c->initiation,
      c->failure, useCompression, batch);
}

void launch_defgrad_with(vpd_context *c, bool dmg, int lds, int rounds, const Batch &batch,
                         const Gather &gx) {
  c->fInN = false; // fnew receives the latest F
  const int useCompression = c->failure > 0.0 && std::isfinite(c->failure);
  const Rounds rd = rounds_of(c->n, rounds);
  if (dmg)
    launch_defgrad_t<true>(c, lds, rd, useCompression, batch, gx);
  else
    launch_defgrad_t<false>(c, lds, rd, useCompression, batch, gx);
}

// F of the step being taken: with batched specimens, of those that have not finished.
int launch_defgrad(vpd_context *c) {
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  if (int s = prepare_plans(c))
    return s;
  return timed(c, K_DEFGRAD, [&] {
    launch_defgrad_with(c, dmg, c->ldsDefgrad, c->roundsDefgrad, c->batch, gather_of(c, 0));
  });
}

// Buffers of the compacted learned update (plastic-point list), allocated on first use.
int ensure_list(vpd_context *c) {
  if (!c->list) {
    c->listZeroed = false;
    if (dalloc(c, c->listCount, 2) || dalloc(c, c->list, c->n) ||
        dalloc(c, c->incList, 8 * c->n))
      return VPD_ERR_DEVICE;
  }
  return VPD_OK;
}

// The compacted material update (k_material_screen, then the return map on the listed
// plastic points only; see vpd_kernels.hpp): in the learned mode the list kernel's last
// wavefronts run the network on the listed points without a positive state 5.
// Blocks of up to kGroupedListMax particles run the bisections' plastic list with
// k_material_return_list_g: on lane groups of kListLanes per point while the list is short.
constexpr int64_t kGroupedListMax = 32768;
// Blocks of up to this many particles run the material screen and the list kernels in
// 64-thread workgroups, spread over the CUs where 256-thread ones would stack four
// wavefronts on a few.
constexpr int64_t kSmallScreenMax = 16384;
constexpr int kListLanes = 32;

template <int MODE> int launch_material_compact(vpd_context *c, MaterialArgs &g, int grid) {
  if (int s = ensure_list(c))
    return s;
  CompactArgs cl{c->listCount, c->list, planes(c->incList, c->n), c->n, nullptr, nullptr};
  if (MODE == RM_LEARNED) {
    cl.theta = c->mlp->theta;
    cl.norm = c->mlp->norm;
  }
  // A step captured by vpd_run_specimens finds them zero at every replay: they are cleared
  // before the first, and each replay's force pass clears them for the next.
  const bool clear = !c->listZeroed;
  c->listZeroed = false;
  if (int s = timed(c, K_MATERIAL, [&] {
        if (clear)
          k_zero_count<<<1, 1, 0, c->stream>>>(c->listCount);
        if (c->n <= kSmallScreenMax) { // one wavefront per workgroup: spread over more CUs
          MaterialArgs gs = g;
          gs.nblocks = ceil_div(c->n, 64);
          k_material_screen<MODE>
              <<<(int)(ceil_div(gs.nblocks, 8) * 8), 64, 0, c->stream>>>(gs, cl);
        } else {
          k_material_screen<MODE><<<grid, 256, 0, c->stream>>>(g, cl);
        }
      }))
    return s;
  return timed(c, K_RETURN_MAP, [&] {
    const int wg = c->n <= kSmallScreenMax ? 64 : 256; // small blocks: spread over the CUs
    if constexpr (MODE == RM_BISECTION || MODE == RM_BISECTION_TIGHT) {
      if (c->n <= kGroupedListMax) { // short lists: lane groups per point
        k_material_return_list_g<MODE, kListLanes>
            <<<(int)ceil_div(c->n * kListLanes, wg), wg, 0, c->stream>>>(g, cl);
        return;
      }
    }
    k_material_return_list<MODE><<<(int)ceil_div(c->n, wg), wg, 0, c->stream>>>(g, cl);
  });
}

template <bool DIAG> int launch_material_t(vpd_context *c, MaterialArgs &g) {
  const int grid = (int)(ceil_div(g.nblocks, 8) * 8); // grid8_of(n), or the diagnostics' chunks
  if (c->materialModel == VPD_MATERIAL_ELASTIC) {
    if (DIAG)
      return fail(c, VPD_ERR_UNSUPPORTED, "Run diagnostics need the Johnson-Cook material.");
    return timed(c, K_MATERIAL,
                 [&] { k_material<RM_ELASTIC, false, false><<<grid, 256, 0, c->stream>>>(g); });
  }
  if (c->returnMap == RM_LEARNED && !c->mlp)
    return fail(c, VPD_ERR_STATE, "The learned return map needs a warm-start network.");
  if (!DIAG) {
    switch (c->returnMap) {
    case RM_BISECTION:
      return launch_material_compact<RM_BISECTION>(c, g, grid);
    case RM_BISECTION_TIGHT:
      return launch_material_compact<RM_BISECTION_TIGHT>(c, g, grid);
    case RM_NEWTON:
      return launch_material_compact<RM_NEWTON>(c, g, grid);
    default:
      return launch_material_compact<RM_LEARNED>(c, g, grid);
    }
  }
  // With run diagnostics each workgroup also reduces its chunk's material terms: one pass
  // over every point.
  switch (c->returnMap) {
  case RM_BISECTION:
    return timed(c, K_MATERIAL,
                 [&] { k_material<RM_BISECTION, false, true><<<grid, 256, 0, c->stream>>>(g); });
  case RM_BISECTION_TIGHT:
    return timed(c, K_MATERIAL, [&] {
      k_material<RM_BISECTION_TIGHT, false, true><<<grid, 256, 0, c->stream>>>(g);
    });
  case RM_NEWTON:
    return timed(c, K_MATERIAL,
                 [&] { k_material<RM_NEWTON, false, true><<<grid, 256, 0, c->stream>>>(g); });
  default: {
    if (int s = timed(c, K_MATERIAL,
                      [&] { k_material<RM_LEARNED, true, true><<<grid, 256, 0, c->stream>>>(g); }))
      return s;
    if (int s = timed(c, K_MLP_SEED, [&] {
          k_mlp_seed<<<mlp_grid(c->n), 256, 0, c->stream>>>(
              c->n, c->inc, c->n, c->statev, c->n, c->mlp->theta, c->mlp->norm, c->seed);
        }))
      return s;
    g.seed = c->seed;
    return timed(c, K_RETURN_MAP, [&] {
      k_material_return<RM_LEARNED, true><<<grid, 256, 0, c->stream>>>(g);
    });
  }
  }
}

// The material stage in the context's return-map mode; with `diag`, the run diagnostics'
// material terms are reduced as well (vpd_step with diagnostics enabled).
int launch_material(vpd_context *c, double dt, bool fields, bool diag = false) {
  if (c->fInN) {
    // The last step swapped the F buffers and no F pass has run since: the material stage
    // updates from the latest F, as it did when the step copied F_{n+1} into both.
    HIPTRY(c, hipMemcpyAsync(c->fnew, c->fN, sizeof(double) * 9 * c->n,
                             hipMemcpyDeviceToDevice, c->stream));
    c->fInN = false;
  }
  MaterialArgs g;
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.fnew = planes(c->fnew, c->n);
  g.fN = planes(c->fN, c->n);
  g.copyF = !c->swapF;
  g.fCur = c->swapF ? g.fnew : g.fN;
  g.vN = planes(c->vN, c->n);
  g.rN = planes(c->rN, c->n);
  g.sig = planes(c->sig, c->n);
  g.statev = planes(c->statev, c->n);
  g.kinv = planes(c->kinv, c->n);
  g.amap = pairs(c->amap, c->n);
  g.temp = c->temp;
  g.eint = c->eint;
  g.einel = c->einel;
  g.sigOut = planes(fields ? c->sigOut : nullptr, c->n);
  g.strainOut = planes(fields ? c->strainOut : nullptr, c->n);
  g.incOut = planes(fields || c->returnMap == RM_LEARNED ? c->inc : nullptr, c->n);
  g.iterOut = fields ? c->iters : nullptr;
  g.fallbackOut = fields ? c->fallbacks : nullptr;
  g.seed = nullptr;
  g.jc = c->jc;
  g.elastic = c->elastic;
  g.dt = dt;
  g.rtol = c->rtol;
  g.err = c->derr;
  g.diag = c->diag;
  g.batch = c->batch;
  if (diag && !c->diagOn)
    return fail(c, VPD_ERR_STATE, "Run diagnostics are not enabled.");
  if (diag && c->diag.chunkFirst) // batched specimens: one workgroup per specimen chunk
    g.nblocks = c->diagChunks;
  return diag ? launch_material_t<true>(c, g) : launch_material_t<false>(c, g);
}

template <bool DAMAGE, bool FUSE, bool INTEGRATE, bool DIAG = false>
void launch_force_t(vpd_context *c, const ForceArgs &g, int lds, const Gather &gx) {
  const int window = gx.window;
  ForceArgs h = g;
  h.take = take_arg(gx.take);
  if (!DIAG && window > 0 && gx.mode == 4)
    with_row_lanes(window, [&](auto w) {
      constexpr int R = decltype(w)::value;
      k_force_m<DAMAGE, FUSE, INTEGRATE, R>
          <<<(int)ceil_div(h.n, 256 / R), 256, lds, c->stream>>>(h);
    });
  else if (!DIAG && window > 0 && gx.mode == 3)
    with_group_window(window, [&](auto w) {
      k_force_g<DAMAGE, FUSE, INTEGRATE, decltype(w)::value>
          <<<grid_rounds(h.rd), 256, std::max(lds, group_lds(7, window)), c->stream>>>(
              h, gx.gplan);
    });
  else if (!DIAG && window > 0 && gx.mode > 0)
    with_planned_window(window, [&](auto w) {
      constexpr int W = decltype(w)::value;
      if (gx.mode == 2)
        k_force_p<DAMAGE, FUSE, INTEGRATE, W, true>
            <<<grid_rounds(h.rd), 256, prefetch_lds(lds, 7, window), c->stream>>>(h, gx.plan);
      else
        k_force_p<DAMAGE, FUSE, INTEGRATE, W, false>
            <<<grid_rounds(h.rd), 256, launch_lds(lds, 7, window), c->stream>>>(h, gx.plan);
    });
  else if (!DIAG && window > 0)
    with_window(window, [&](auto w) {
      k_force_w<DAMAGE, FUSE, INTEGRATE, decltype(w)::value>
          <<<grid_rounds(h.rd), 256, launch_lds(lds, 7, window), c->stream>>>(h);
    });
  else
    k_force<DAMAGE, FUSE, INTEGRATE, DIAG><<<grid_rounds(h.rd), 256, lds, c->stream>>>(h);
}

template <bool DIAG>
void launch_force_integrating(vpd_context *c, const ForceArgs &g, bool dmg, bool fuse, int lds,
                              const Gather &gx) {
  if (dmg && fuse)
    launch_force_t<true, true, true, DIAG>(c, g, lds, gx);
  else if (dmg)
    launch_force_t<true, false, true, DIAG>(c, g, lds, gx);
  else if (fuse)
    launch_force_t<false, true, true, DIAG>(c, g, lds, gx);
  else
    launch_force_t<false, false, true, DIAG>(c, g, lds, gx);
}

// The force pass in one of its forms: with or without damage flags, the next step's predictor
// fused in, the integrator (INTEGRATE false leaves only the force density), the run
// diagnostics' force terms (integrating forms only, direct gathers), direct or windowed
// gathers (window > 0).
void launch_force_with(vpd_context *c, const ForceArgs &g, bool dmg, bool fuse, bool integrate,
                       int lds, bool diag = false, const Gather &gx = Gather{}) {
  if (!integrate) {
    if (dmg)
      launch_force_t<true, false, false>(c, g, lds, gx);
    else
      launch_force_t<false, false, false>(c, g, lds, gx);
  } else if (diag) {
    launch_force_integrating<true>(c, g, dmg, fuse, lds, Gather{});
  } else {
    launch_force_integrating<false>(c, g, dmg, fuse, lds, gx);
  }
}

int launch_force(vpd_context *c, double dt, const double ramp[3], bool fuse,
                 const double *rampNext, bool fields, bool diag = false) {
  ForceArgs g;
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.x = planes(c->x, c->n);
  g.u = upairs(c);
  g.v = planes(c->v, c->n);
  g.a = planes(c->a, c->n);
  g.amap = pairs(c->amap, c->n);
  g.xv = pairs(c->xv, c->n);
  g.forceOut = planes(fields ? c->force : nullptr, c->n);
  g.vol = c->vol;
  g.bonds = sell_of(c);
  g.broken = c->broken;
  g.anyBroken = c->anyBroken;
  g.codes = c->codes;
  g.dt = dt;
  g.density = c->density;
  g.rampNow = ramp_of(ramp);
  g.rampNext = fuse ? ramp_of(rampNext) : Ramp{0, 0, 0};
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  g.rd = rounds_of(c->n, c->roundsForce);
  if (diag && c->diag.chunkFirst) { // batched specimens: one workgroup per specimen chunk
    g.rd.rounds = 1;
    g.rd.blocksPerRound = c->diagChunks;
  }
  g.diag = c->diag;
  g.batch = c->batch;
  if (int s = prepare_plans(c))
    return s;
  g.zeroCount = c->listCount; // the pass clears the material lists for the next update
  g.stepNext = c->stepNextDev; // captured steps: the next step's counter slot
  if (int s = timed(c, K_FORCE, [&] {
        launch_force_with(c, g, dmg, fuse, true, c->ldsForce, diag, gather_of(c, 1));
      }))
    return s;
  c->listZeroed = c->listCount != nullptr; // (also in a captured step: each replay clears)
  return VPD_OK;
}

int require_block(vpd_context *c) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  if (c->n <= 0)
    return fail(c, VPD_ERR_STATE, "Invalid or incomplete PD particle block.");
  return VPD_OK;
}

// The scalar-time entry points drive a block that is one specimen.
int require_single(vpd_context *c) {
  if (c && c->batch.id)
    return fail(c, VPD_ERR_STATE,
                "The particle block holds batched specimens; step it with vpd_step_specimens.");
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
  if (cfg->material_model != VPD_MATERIAL_JOHNSON_COOK &&
      cfg->material_model != VPD_MATERIAL_ELASTIC)
    return fail(nullptr, VPD_ERR_UNSUPPORTED, "Unsupported material model.");
  const bool elastic = cfg->material_model == VPD_MATERIAL_ELASTIC;
  if (material && !elastic && (cfg->nprops < 14 || cfg->nstatev < 6))
    return fail(nullptr, VPD_ERR_ARGUMENT, "JohnsonCook VUMAT requires 14 props and 6 states");
  if (material && elastic && (cfg->nprops < 2 || cfg->nstatev < 0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "The elastic VUMAT requires 2 props (E, nu).");
  if (cfg->damage_mode != VPD_DAMAGE_NONE && cfg->damage_mode != VPD_DAMAGE_CRITICAL_STRETCH)
    return fail(nullptr, VPD_ERR_UNSUPPORTED, "Unsupported damage mode.");
  if (cfg->damage_mode == VPD_DAMAGE_CRITICAL_STRETCH && !(cfg->damage_initiation > 0.0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "critical_stretch requires initiation > 0.");
  if (cfg->return_map < VPD_RETURN_MAP_BISECTION || cfg->return_map > VPD_RETURN_MAP_LEARNED)
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
  c->materialModel = cfg->material_model;
  if (material && elastic)
    c->elastic = make_elastic(c->props[0], c->props[1]);
  else if (material)
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
  dfree(c->xv);
  dfree(c->sigOut);
  dfree(c->strainOut);
  dfree(c->force);
  dfree(c->dmgIndex);
  dfree(c->iters);
  dfree(c->fallbacks);
  dfree(c->inc);
  dfree(c->seed);
  dfree(c->listCount);
  dfree(c->list);
  dfree(c->incList);
  dfree(c->broken);
  dfree(c->anyBroken);
  dfree(c->colS);
  dfree(c->sliceBase);
  for (auto &p : c->plan) {
    dfree(p.first);
    dfree(p.start);
    dfree(p.rec);
  }
  for (auto &p : c->gplan) {
    dfree(p.first);
    dfree(p.meta);
    dfree(p.off);
  }
  dfree(c->diagPartials);
  dfree(c->diagRecords);
  dfree(c->diagChunkFirst);
  dfree(c->diagChunkEnd);
  dfree(c->diagSpecChunk);
  dfree(c->specBand);
  dfree(c->specBandOff);
  dfree(c->diagW);
  dfree(c->diagDstrain);
  dfree(c->diagLat);
  dfree(c->specId);
  dfree(c->specDt);
  dfree(c->specVel);
  dfree(c->specLast);
  dfree(c->stepDev);
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
  c->diagOn = false; // sized for the previous block
  c->batch = Batch{};
  c->specCount = 0;
  const int64_t sz[] = {3 * n, n, 9 * n, 4 * n, 3 * n, 3 * n, 9 * n, 9 * n,
                        9 * n, 6 * n, (int64_t)c->nstatev * n, n, n, n, 9 * n, 10 * n, 4 * n};
  double **bufs[] = {&c->x,  &c->vol, &c->kinv, &c->u,    &c->v,   &c->a,
                     &c->fN, &c->vN,  &c->rN,   &c->sig,  &c->statev, &c->temp,
                     &c->eint, &c->einel, &c->fnew, &c->amap, &c->xv};
  for (size_t k = 0; k < sizeof(sz) / sizeof(sz[0]); ++k)
    if (dalloc(c, *bufs[k], sz[k]))
      return VPD_ERR_DEVICE;
  if (dalloc(c, c->codes, n) || dalloc(c, c->inc, 6 * n) || dalloc(c, c->seed, n))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemsetAsync(c->codes, 0, n, c->stream));
  HIPTRY(c, hipMemsetAsync(c->u, 0, sizeof(double) * 4 * n, c->stream));
  dfree(c->sigOut);
  dfree(c->strainOut);
  dfree(c->force);
  dfree(c->dmgIndex);
  dfree(c->iters);
  dfree(c->fallbacks);
  dfree(c->listCount); // sized for the previous block
  dfree(c->list);
  dfree(c->incList);
  c->haveNeighbors = c->haveShape = c->haveMaterial = c->haveDynamics = false;
  c->haveFnew = c->haveFields = c->haveForce = c->predicted = false;
  if (int s = put_planes(c, c->x, x, 3))
    return s;
  HIPTRY(c, hipMemcpyAsync(c->vol, vol, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  k_pack_xv<<<grid_of(n), 256, 0, c->stream>>>(n, planes(c->x, n), c->vol, pairs(c->xv, n));
  HIPTRY(c, hipGetLastError());
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
  // Every particle, whatever step a batched specimen has reached.
  if (int s = prepare_plans(c))
    return s;
  if (int s = timed(c, K_DEFGRAD, [&] {
        launch_defgrad_with(c, false, c->ldsDefgrad, c->roundsDefgrad, Batch{},
                            gather_of(c, 0));
      }))
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
  if (c->anyBroken)
    HIPTRY(c, hipMemsetAsync(c->anyBroken, 0, std::max<int64_t>(1, c->n), c->stream));
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
    std::vector<uint8_t> any(std::max<int64_t>(1, c->n), 0);
    for (int64_t i = 0; i < c->n; ++i)
      for (int64_t p = c->hostRowptr[i]; p < c->hostRowptr[i + 1]; ++p) {
        const int64_t q = sell_slot(c, i, p);
        if (std::min(std::max(d[p], 0.0), 1.0) >= 1.0) {
          bits[q >> 5] |= 1u << (q & 31);
          any[i] = 1;
        } else if (d[p] > 0.0) {
          return fail(c, VPD_ERR_UNSUPPORTED, "Partial bond damage is not representable.");
        }
      }
    HIPTRY(c, hipMemcpyAsync(c->broken, bits.data(), sizeof(uint32_t) * bits.size(),
                             hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, hipMemcpyAsync(c->anyBroken, any.data(), any.size(), hipMemcpyHostToDevice,
                             c->stream));
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
  if (field == VPD_FIELD_DEFGRAD)
    c->fInN = false; // the caller's F goes to fnew, the state's F_n stays in fN
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
  if (field == VPD_FIELD_DISPLACEMENT)
    return put_u(c, static_cast<const double *>(host));
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
  if (field == VPD_FIELD_STRAIN_INCREMENT) {
    if (!c->haveFields)
      return fail(c, VPD_ERR_STATE, "Field is not available: run a step with want_fields.");
    return get_planes(c, c->inc, static_cast<double *>(host), 6);
  }
  if (field == VPD_FIELD_RETURN_MAP_FALLBACKS) {
    if (!c->fallbacks || !c->haveFields)
      return fail(c, VPD_ERR_STATE, "Field is not available: run a step with want_fields.");
    HIPTRY(c, hipMemcpyAsync(host, c->fallbacks, sizeof(int) * c->n, hipMemcpyDeviceToHost,
                             c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
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
  if (field == VPD_FIELD_DISPLACEMENT)
    return get_u(c, static_cast<double *>(host));
  return get_planes(c, src, static_cast<double *>(host), components(c, field));
}

int vpd_predict(vpd_context *c, double dt, const double ramp[3]) {
  if (int s = require_block(c))
    return s;
  if (int s = require_single(c))
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
  if (int s = require_single(c))
    return s;
  if (!ramp)
    return fail(c, VPD_ERR_ARGUMENT, "Null ramp.");
  return timed(c, K_CONSTRAIN, [&] {
    k_constrain<<<grid_of(c->n), 256, 0, c->stream>>>(c->n, upairs(c),
                                                      planes(c->v, c->n), planes(c->a, c->n),
                                                      c->codes, ramp_of(ramp), c->batch);
  });
}

int vpd_compute_deformation_gradient(vpd_context *c) {
  if (int s = require_block(c))
    return s;
  if (int s = require_single(c))
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
  if (int s = require_single(c))
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
  if (int s = require_single(c))
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
  if (int s = require_single(c))
    return s;
  if (!g || !(g->dt > 0.0))
    return fail(c, VPD_ERR_ARGUMENT, "Invalid explicit dynamics predictor input.");
  if (g->want_fields)
    if (int s = ensure_outputs(c))
      return s;
  if (int s = prepare_plans(c)) // before the step changes any state
    return s;
  const StepTiming stepTiming(c);
  if (!c->predicted)
    if (int s = launch_predict(c, g->dt, g->ramp))
      return s;
  if (c->diagOn && c->diagCount >= c->diagCapacity)
    return fail(c, VPD_ERR_STATE,
                "The diagnostics record buffer is full; fetch the records first.");
  if (int s = launch_defgrad(c))
    return s;
  // One specimen, eager, no run diagnostics, the JC material: F_{n+1} stays in fnew and the
  // buffers swap instead of the material stage copying it into fN.
  c->swapF = !c->batch.id && !c->capturing && !c->diagOn &&
             c->materialModel != VPD_MATERIAL_ELASTIC;
  const int ms = launch_material(c, g->dt, g->want_fields != 0, c->diagOn);
  if (c->swapF && !ms) {
    std::swap(c->fN, c->fnew);
    c->fInN = true;
  }
  c->swapF = false;
  if (ms)
    return ms;
  if (int s = launch_force(c, g->dt, g->ramp, g->fuse_next != 0, g->ramp_next,
                           g->want_fields != 0, c->diagOn))
    return s;
  if (c->diagOn) {
    k_diag_reduce<<<1, 256, 0, c->stream>>>(c->diagPartials, ceil_div(c->n, 256),
                                            c->diagRecords + (int64_t)c->diagCount * DIAG_STRIDE);
    HIPTRY(c, hipGetLastError());
    ++c->diagCount;
  }
  c->predicted = g->fuse_next != 0;
  c->haveFnew = true;
  c->haveFields = g->want_fields != 0;
  c->haveForce = g->want_fields != 0;
  return VPD_OK;
}

int vpd_set_specimens(vpd_context *c, int count, const int64_t *first, const double *dt,
                      const double *velocity, const int32_t *lastStep) {
  if (int s = require_ready(c))
    return s;
  if (count == 0) { // back to one specimen
    if (c->diagOn)
      return fail(c, VPD_ERR_STATE,
                  "Disable run diagnostics before leaving the specimen table.");
    c->batch = Batch{};
    c->specCount = 0;
    c->predicted = false;
    return VPD_OK;
  }
  if (c->diagOn)
    return fail(c, VPD_ERR_STATE,
                "Set the specimen table before enabling run diagnostics (or disable them).");
  if (count < 0 || count > 65536 || !first || !dt || !velocity || !lastStep)
    return fail(c, VPD_ERR_ARGUMENT, "Invalid specimen table.");
  if (first[0] != 0 || first[count] != c->n)
    return fail(c, VPD_ERR_ARGUMENT, "Specimen ranges must cover the particle block.");
  for (int s = 0; s < count; ++s) {
    if (first[s + 1] <= first[s])
      return fail(c, VPD_ERR_ARGUMENT, "Specimen ranges must be ascending and non-empty.");
    if (!(dt[s] > 0.0) || !std::isfinite(dt[s]) || !std::isfinite(velocity[s]))
      return fail(c, VPD_ERR_ARGUMENT, "Specimen time steps must be positive and finite.");
    if (lastStep[s] < 1)
      return fail(c, VPD_ERR_ARGUMENT, "Every specimen takes at least one step.");
  }
  // A bond between two specimens would couple them: refuse it.
  std::vector<int> col(std::max<int64_t>(1, c->nnz));
  if (c->nnz)
    HIPTRY(c, hipMemcpy(col.data(), c->col, sizeof(int) * c->nnz, hipMemcpyDeviceToHost));
  std::vector<uint16_t> id(c->n);
  for (int s = 0; s < count; ++s)
    for (int64_t i = first[s]; i < first[s + 1]; ++i)
      id[i] = (uint16_t)s;
  for (int64_t i = 0; i < c->n; ++i)
    for (int p = c->hostRowptr[i]; p < c->hostRowptr[i + 1]; ++p)
      if (id[col[p]] != id[i])
        return fail(c, VPD_ERR_ARGUMENT,
                    "Particles " + std::to_string(i) + " and " + std::to_string(col[p]) +
                        " of different specimens are bonded; keep specimens more than a "
                        "horizon apart.");
  if (dalloc(c, c->specId, c->n) || dalloc(c, c->specDt, count) ||
      dalloc(c, c->specVel, count) || dalloc(c, c->specLast, count) ||
      (!c->stepDev && dalloc(c, c->stepDev, 2)))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemcpy(c->specId, id.data(), sizeof(uint16_t) * c->n, hipMemcpyHostToDevice));
  HIPTRY(c, hipMemcpy(c->specDt, dt, sizeof(double) * count, hipMemcpyHostToDevice));
  HIPTRY(c, hipMemcpy(c->specVel, velocity, sizeof(double) * count, hipMemcpyHostToDevice));
  HIPTRY(c, hipMemcpy(c->specLast, lastStep, sizeof(int32_t) * count, hipMemcpyHostToDevice));
  c->specCount = count;
  c->specFirst.assign(first, first + count + 1);
  c->specLastHost.assign(lastStep, lastStep + count);
  c->batch = Batch{c->specId, c->specDt, c->specVel, c->specLast, 0, nullptr, nullptr, nullptr};
  c->predicted = false;
  // Each specimen's loaded band at t = 0 (the single-specimen host's initial constraints).
  if (int s = timed(c, K_CONSTRAIN, [&] {
        k_constrain<<<grid_of(c->n), 256, 0, c->stream>>>(
            c->n, upairs(c), planes(c->v, c->n), planes(c->a, c->n), c->codes,
            Ramp{0, 0, 0}, c->batch);
      }))
    return s;
  return check_device(c);
}

int vpd_set_specimen_bands(vpd_context *c, const int64_t *offsets, const double *table) {
  if (int s = require_ready(c))
    return s;
  if (!c->batch.id)
    return fail(c, VPD_ERR_STATE, "No specimen table; call vpd_set_specimens first.");
  if (c->batch.step != 0 || c->predicted)
    return fail(c, VPD_ERR_STATE, "Set the specimen bands before the first step.");
  if (offsets) {
    const int count = c->specCount;
    if (offsets[0] != 0)
      return fail(c, VPD_ERR_ARGUMENT, "Band table offsets must start at 0.");
    for (int sp = 0; sp < count; ++sp) {
      const int64_t rows = offsets[sp + 1] - offsets[sp];
      if (rows != 0 && rows != (int64_t)c->specLastHost[sp] + 1)
        return fail(c, VPD_ERR_ARGUMENT,
                    "A specimen's band table holds one row per step 0.." +
                        std::to_string(c->specLastHost[sp]) + " (or none).");
    }
    const int64_t rows = offsets[count];
    if (rows > 0 && !table)
      return fail(c, VPD_ERR_ARGUMENT, "Null band table.");
    for (int64_t k = 0; k < 3 * rows; ++k)
      if (!std::isfinite(table[k]))
        return fail(c, VPD_ERR_ARGUMENT, "Band table values must be finite.");
    if (dalloc(c, c->specBand, std::max<int64_t>(1, 3 * rows)) ||
        dalloc(c, c->specBandOff, count + 1))
      return VPD_ERR_DEVICE;
    if (rows > 0)
      HIPTRY(c, hipMemcpy(c->specBand, table, sizeof(double) * 3 * rows, hipMemcpyHostToDevice));
    HIPTRY(c, hipMemcpy(c->specBandOff, offsets, sizeof(int64_t) * (count + 1),
                        hipMemcpyHostToDevice));
    c->batch.band = c->specBand;
    c->batch.bandOff = c->specBandOff;
  } else {
    c->batch.band = nullptr;
    c->batch.bandOff = nullptr;
  }
  // Every specimen's band at t = 0 again (the single-specimen host's initial constraints).
  if (int s = timed(c, K_CONSTRAIN, [&] {
        k_constrain<<<grid_of(c->n), 256, 0, c->stream>>>(
            c->n, upairs(c), planes(c->v, c->n), planes(c->a, c->n), c->codes,
            Ramp{0, 0, 0}, c->batch);
      }))
    return s;
  return check_device(c);
}

int vpd_step_specimens(vpd_context *c, int32_t step, int fuseNext, int wantFields) {
  if (int s = require_ready(c))
    return s;
  if (!c->batch.id)
    return fail(c, VPD_ERR_STATE, "No specimen table; call vpd_set_specimens first.");
  if (step < 1)
    return fail(c, VPD_ERR_ARGUMENT, "Specimen steps count from 1.");
  if (c->predicted && step != c->batch.step + 1)
    return fail(c, VPD_ERR_STATE, "The fused prediction was made for step " +
                                      std::to_string(c->batch.step + 1) + ".");
  if (wantFields)
    if (int s = ensure_outputs(c))
      return s;
  if (int s = prepare_plans(c)) // before the step changes any state
    return s;
  const StepTiming stepTiming(c);
  c->batch.step = step;
  // Scalar arguments are unused: every particle takes its specimen's dt and velocity band.
  const double none[3] = {0.0, 0.0, 0.0};
  if (!c->predicted)
    if (int s = launch_predict(c, 1.0, none))
      return s;
  if (c->diagOn && c->diagCount >= c->diagCapacity)
    return fail(c, VPD_ERR_STATE,
                "The diagnostics record buffer is full; fetch the records first.");
  if (int s = launch_defgrad(c))
    return s;
  if (int s = launch_material(c, 1.0, wantFields != 0, c->diagOn))
    return s;
  if (int s = launch_force(c, 1.0, none, fuseNext != 0, none, wantFields != 0, c->diagOn))
    return s;
  if (c->diagOn) { // one record per specimen
    k_diag_reduce_specimens<<<c->specCount, 256, 0, c->stream>>>(
        c->diagPartials, c->diagSpecChunk,
        c->diagRecords + (int64_t)c->diagCount * c->diagPerStep * DIAG_STRIDE);
    HIPTRY(c, hipGetLastError());
    ++c->diagCount;
  }
  c->predicted = fuseNext != 0;
  c->haveFnew = true;
  c->haveFields = wantFields != 0;
  c->haveForce = wantFields != 0;
  return VPD_OK;
}

int vpd_run_specimens(vpd_context *c, int32_t firstStep, int32_t count) {
  if (int s = require_ready(c))
    return s;
  if (!c->batch.id)
    return fail(c, VPD_ERR_STATE, "No specimen table; call vpd_set_specimens first.");
  if (firstStep < 1 || count < 1)
    return fail(c, VPD_ERR_ARGUMENT, "Specimen steps count from 1; count must be positive.");
  if (c->diagOn)
    return fail(c, VPD_ERR_UNSUPPORTED,
                "Graph-replayed steps take no run diagnostics; use vpd_step_specimens.");
  if (c->predicted && firstStep != c->batch.step + 1)
    return fail(c, VPD_ERR_STATE, "The fused prediction was made for step " +
                                      std::to_string(c->batch.step + 1) + ".");
  if (c->returnMap == RM_LEARNED && !c->mlp)
    return fail(c, VPD_ERR_STATE, "The learned return map needs a warm-start network.");
  if (c->materialModel != VPD_MATERIAL_ELASTIC)
    if (int s = ensure_list(c)) // no allocation may happen while the step is captured
      return s;
  if (int s = prepare_plans(c)) // nor a plan build
    return s;
  c->batch.step = firstStep;
  const double none[3] = {0.0, 0.0, 0.0};
  if (!c->predicted)
    if (int s = launch_predict(c, 1.0, none))
      return s;
  k_set_step<<<1, 64, 0, c->stream>>>(c->stepDev, firstStep);
  if (c->listCount && !c->listZeroed) // the first replay's list lengths (see below)
    k_zero_count<<<1, 1, 0, c->stream>>>(c->listCount);
  c->listZeroed = c->listCount != nullptr;
  HIPTRY(c, hipGetLastError());
  // Two fused steps captured once: the first reads its number from counter slot 0 and its
  // force pass sets slot 1, the second reads slot 1 and sets slot 0 for the next replay, so
  // the graph needs no kernel of its own to move the counter. An odd last step runs eagerly.
  const bool timing = c->timing;
  const Batch eager = c->batch;
  c->timing = false;
  hipGraphExec_t exec = nullptr;
  const int pairs = count / 2;
  if (pairs > 0) {
    hipGraph_t graph = nullptr;
    HIPTRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int s = VPD_OK;
    for (int h = 0; h < 2 && !s; ++h) {
      c->batch.stepDev = c->stepDev + h;
      c->stepNextDev = c->stepDev + (1 - h);
      s = launch_defgrad(c);
      if (!s)
        s = launch_material(c, 1.0, false);
      if (!s)
        s = launch_force(c, 1.0, none, true, none, false);
    }
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    c->capturing = false;
    c->stepNextDev = nullptr;
    c->batch = eager;
    if (s) {
      if (ec == hipSuccess && graph)
        (void)hipGraphDestroy(graph);
      c->listZeroed = false;
      c->timing = timing;
      return s;
    }
    HIPTRY(c, ec);
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPTRY(c, ei);
  }
  hipError_t el = hipSuccess;
  for (int k = 0; k < pairs && el == hipSuccess; ++k)
    el = hipGraphLaunch(exec, c->stream);
  int s = VPD_OK;
  if (el == hipSuccess && count % 2) { // the odd last step, eagerly, as the captured ones
    c->batch.step = firstStep + count - 1;
    s = launch_defgrad(c);
    if (!s)
      s = launch_material(c, 1.0, false);
    if (!s)
      s = launch_force(c, 1.0, none, true, none, false);
    c->batch = eager;
  }
  c->timing = timing;
  const hipError_t es = hipStreamSynchronize(c->stream);
  if (exec)
    (void)hipGraphExecDestroy(exec);
  c->listZeroed = false;
  HIPTRY(c, el);
  HIPTRY(c, es);
  if (s)
    return s;
  c->batch.step = firstStep + count - 1;
  c->predicted = true; // every specimen past its last step, or predicted for the next one
  c->haveFnew = true;
  c->haveFields = c->haveForce = false;
  return check_device(c);
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
                                                planes(c->fInN ? c->fN : c->fnew, c->n),
                                                planes(c->kinv, c->n),
                                                pairs(c->amap, c->n), c->derr);
  HIPTRY(c, hipGetLastError());
  if (int s = check_device(c))
    return s;
  ForceArgs g{};
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.x = planes(c->x, c->n);
  g.amap = pairs(c->amap, c->n);
  g.xv = pairs(c->xv, c->n);
  g.forceOut = planes(c->force, c->n);
  g.vol = c->vol;
  g.bonds = sell_of(c);
  g.broken = c->broken;
  g.anyBroken = c->anyBroken;
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  g.rd = rounds_of(c->n, c->roundsForce);
  if (int s = timed(c, K_FORCE, [&] {
        launch_force_with(c, g, dmg, false, false, c->ldsForce);
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
                                                      upairs(c), sell_of(c), c->broken,
                                                      c->anyBroken, c->initiation, c->failure,
                                                      useCompression);
  HIPTRY(c, hipGetLastError());
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return VPD_OK;
}

int vpd_enable_diagnostics(vpd_context *c, const vpd_diag_config *cfg) {
  if (int s = require_block(c))
    return s;
  if (!cfg) {
    c->diagOn = false;
    c->diagCount = 0;
    return VPD_OK;
  }
  if (c->materialModel != VPD_MATERIAL_JOHNSON_COOK)
    return fail(c, VPD_ERR_UNSUPPORTED, "Run diagnostics need the Johnson-Cook material.");
  const int ax = cfg->load_axis, lat = cfg->lateral_axis;
  if (ax < 0 || ax > 2 || lat < 0 || lat > 2)
    return fail(c, VPD_ERR_ARGUMENT, "Diagnostics axes must lie in [0, 2].");
  if (cfg->capacity < 1)
    return fail(c, VPD_ERR_ARGUMENT, "Diagnostics need room for at least one step record.");
  // The material kernels hold the Johnson-Cook model's 6 state variables in registers;
  // entries beyond them never change from their initial values.
  const int held = std::min(c->nstatev, 6);
  if (cfg->eqps_index < -1 || cfg->eqps_index >= held)
    return fail(c, VPD_ERR_ARGUMENT,
                "equivalentPlasticStrainIndex must name one of the first 6 state variables.");
  if (cfg->indicator_count < 0 || cfg->indicator_count > 8)
    return fail(c, VPD_ERR_ARGUMENT, "At most 8 internal-indicator state indices.");
  for (int k = 0; k < cfg->indicator_count; ++k)
    if (cfg->indicator_indices[k] < 0 || cfg->indicator_indices[k] >= held)
      return fail(c, VPD_ERR_ARGUMENT,
                  "internalIndicator indices must name the first 6 state variables.");
  // With batched specimens every chunk lies in one specimen and starts where the specimen
  // alone would start it, so each specimen's record equals its single run's bit for bit.
  const int perStep = c->batch.id ? c->specCount : 1;
  std::vector<int64_t> chunkFirst, chunkEnd, specChunk;
  if (c->batch.id) {
    for (int sp = 0; sp < c->specCount; ++sp) {
      specChunk.push_back((int64_t)chunkFirst.size());
      for (int64_t i = c->specFirst[sp]; i < c->specFirst[sp + 1]; i += 256) {
        chunkFirst.push_back(i);
        chunkEnd.push_back(c->specFirst[sp + 1]);
      }
    }
    specChunk.push_back((int64_t)chunkFirst.size());
  }
  const int64_t chunks = c->batch.id ? (int64_t)chunkFirst.size() : ceil_div(c->n, 256);
  if (dalloc(c, c->diagPartials, chunks * DIAG_STRIDE) ||
      dalloc(c, c->diagRecords, (int64_t)cfg->capacity * perStep * DIAG_STRIDE) ||
      dalloc(c, c->diagDstrain, 6 * c->n))
    return VPD_ERR_DEVICE;
  DiagArgs d{};
  d.partials = c->diagPartials;
  if (c->batch.id) {
    if (dalloc(c, c->diagChunkFirst, chunks) || dalloc(c, c->diagChunkEnd, chunks) ||
        dalloc(c, c->diagSpecChunk, c->specCount + 1))
      return VPD_ERR_DEVICE;
    HIPTRY(c, hipMemcpyAsync(c->diagChunkFirst, chunkFirst.data(), sizeof(int64_t) * chunks,
                             hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, hipMemcpyAsync(c->diagChunkEnd, chunkEnd.data(), sizeof(int64_t) * chunks,
                             hipMemcpyHostToDevice, c->stream));
    HIPTRY(c, hipMemcpyAsync(c->diagSpecChunk, specChunk.data(),
                             sizeof(int64_t) * specChunk.size(), hipMemcpyHostToDevice,
                             c->stream));
    d.chunkFirst = c->diagChunkFirst;
    d.chunkEnd = c->diagChunkEnd;
  }
  d.chunks = chunks;
  d.dstrain = planes(c->diagDstrain, c->n);
  d.vol = c->vol;
  if (cfg->observation_weight) {
    if (dalloc(c, c->diagW, c->n))
      return VPD_ERR_DEVICE;
    HIPTRY(c, hipMemcpyAsync(c->diagW, cfg->observation_weight, sizeof(double) * c->n,
                             hipMemcpyHostToDevice, c->stream));
    d.obsW = c->diagW;
  }
  if (cfg->lateral_sets) {
    if (dalloc(c, c->diagLat, c->n))
      return VPD_ERR_DEVICE;
    HIPTRY(c, hipMemcpyAsync(c->diagLat, cfg->lateral_sets, c->n, hipMemcpyHostToDevice,
                             c->stream));
    d.latSet = c->diagLat;
  }
  d.axis = ax;
  d.lateral = lat;
  d.eqpsIndex = cfg->eqps_index;
  d.nstatev = held;
  d.indCount = cfg->indicator_count;
  d.indMin = cfg->indicator_min != 0;
  d.indOneMinus = cfg->indicator_one_minus != 0;
  for (int k = 0; k < 8; ++k)
    d.ind[k] = k < cfg->indicator_count ? cfg->indicator_indices[k] : 0;
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->diag = d;
  c->diagChunks = chunks;
  c->diagPerStep = perStep;
  c->diagCapacity = cfg->capacity;
  c->diagCount = 0;
  c->diagOn = true;
  return VPD_OK;
}

int vpd_fetch_diagnostics(vpd_context *c, double *records, int capacity, int *count) {
  if (!c || !count || (capacity > 0 && !records))
    return fail(c, VPD_ERR_ARGUMENT, "Null argument.");
  if (!c->diagOn)
    return fail(c, VPD_ERR_STATE, "Run diagnostics are not enabled.");
  if (capacity < c->diagCount)
    return fail(c, VPD_ERR_ARGUMENT, "The record array is smaller than the pending records.");
  if (int s = check_device(c))
    return s;
  if (c->diagCount > 0)
    HIPTRY(c, hipMemcpyAsync(records, c->diagRecords,
                             sizeof(double) * DIAG_STRIDE * (size_t)c->diagPerStep *
                                 (size_t)c->diagCount,
                             hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  *count = c->diagCount;
  c->diagCount = 0;
  return VPD_OK;
}

int vpd_set_tuning(vpd_context *c, int knob, int value) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  if (knob == VPD_TUNE_DEFGRAD_LDS || knob == VPD_TUNE_FORCE_LDS) {
    if (value < 0 || value > 65536)
      return fail(c, VPD_ERR_ARGUMENT, "LDS reservation must lie in [0, 65536] bytes.");
    (knob == VPD_TUNE_DEFGRAD_LDS ? c->ldsDefgrad : c->ldsForce) = value;
  } else if (knob == VPD_TUNE_DEFGRAD_ROUNDS || knob == VPD_TUNE_FORCE_ROUNDS) {
    if (value < 1 || value > 1024)
      return fail(c, VPD_ERR_ARGUMENT, "Sweeps per bond pass must lie in [1, 1024].");
    (knob == VPD_TUNE_DEFGRAD_ROUNDS ? c->roundsDefgrad : c->roundsForce) = value;
  } else if (knob == VPD_TUNE_DEFGRAD_WINDOW || knob == VPD_TUNE_FORCE_WINDOW) {
    if (value != 0 && !row_lanes_supported(value) && !window_supported(value) &&
        !group_window_supported(value))
      return fail(c, VPD_ERR_ARGUMENT,
                  "Gather window must be 0 (direct) or one of 64, 80, 96, 128, 160 particles "
                  "(256, 384, 480, 512, 640 with group windows; 16, 32, 64 lanes per particle "
                  "with shared rows).");
    (knob == VPD_TUNE_DEFGRAD_WINDOW ? c->winDefgrad : c->winForce) = value;
  } else if (knob == VPD_TUNE_DEFGRAD_WINDOW_BONDS || knob == VPD_TUNE_FORCE_WINDOW_BONDS) {
    if (value < 0 || value > kMaxTake)
      return fail(c, VPD_ERR_ARGUMENT,
                  "Bonds per lane and window must lie in [0, 1024] (0: no limit).");
    (knob == VPD_TUNE_DEFGRAD_WINDOW_BONDS ? c->takeDefgrad : c->takeForce) = value;
  } else if (knob == VPD_TUNE_DEFGRAD_WINDOW_PLAN || knob == VPD_TUNE_FORCE_WINDOW_PLAN) {
    if (value < 0 || value > 4)
      return fail(c, VPD_ERR_ARGUMENT, "Window form must be 0 (walk), 1 (planned), 2 (planned "
                                       "with prefetch), 3 (group windows) or 4 (shared rows).");
    (knob == VPD_TUNE_DEFGRAD_WINDOW_PLAN ? c->modeDefgrad : c->modeForce) = value;
  } else {
    return fail(c, VPD_ERR_ARGUMENT, "Unknown tuning knob.");
  }
  return VPD_OK;
}

int vpd_autotune_candidates(vpd_context *c, int pass, int max, int *knobs, float *ms,
                            int *count) {
  if (!c || !count || (max > 0 && !knobs))
    return fail(c, VPD_ERR_ARGUMENT, "Null argument.");
  if (pass != 0 && pass != 1)
    return fail(c, VPD_ERR_ARGUMENT, "Bond pass must be 0 (F) or 1 (force).");
  const auto &top = c->tunedTop[pass];
  const int k = std::max(0, std::min(max, (int)top.size()));
  for (int r = 0; r < k; ++r) {
    int *t = knobs + (int64_t)r * VPD_TUNE_COUNT;
    for (int knob = 0; knob < VPD_TUNE_COUNT; ++knob)
      if (int s = vpd_get_tuning(c, knob, t + knob))
        return s;
    const vpd_context::TunedSetting &x = top[r];
    t[pass == 0 ? VPD_TUNE_DEFGRAD_LDS : VPD_TUNE_FORCE_LDS] = x.lds;
    t[pass == 0 ? VPD_TUNE_DEFGRAD_ROUNDS : VPD_TUNE_FORCE_ROUNDS] = x.rounds;
    t[pass == 0 ? VPD_TUNE_DEFGRAD_WINDOW : VPD_TUNE_FORCE_WINDOW] = x.window;
    t[pass == 0 ? VPD_TUNE_DEFGRAD_WINDOW_BONDS : VPD_TUNE_FORCE_WINDOW_BONDS] = x.take;
    t[pass == 0 ? VPD_TUNE_DEFGRAD_WINDOW_PLAN : VPD_TUNE_FORCE_WINDOW_PLAN] = x.mode;
    if (ms)
      ms[r] = x.ms;
  }
  *count = k;
  return VPD_OK;
}

int vpd_get_tuning(vpd_context *c, int knob, int *value) {
  if (!c || !value)
    return fail(c, VPD_ERR_ARGUMENT, "Null argument.");
  switch (knob) {
  case VPD_TUNE_DEFGRAD_LDS:
    *value = c->ldsDefgrad;
    return VPD_OK;
  case VPD_TUNE_FORCE_LDS:
    *value = c->ldsForce;
    return VPD_OK;
  case VPD_TUNE_DEFGRAD_ROUNDS:
    *value = c->roundsDefgrad;
    return VPD_OK;
  case VPD_TUNE_FORCE_ROUNDS:
    *value = c->roundsForce;
    return VPD_OK;
  case VPD_TUNE_DEFGRAD_WINDOW:
    *value = c->winDefgrad;
    return VPD_OK;
  case VPD_TUNE_FORCE_WINDOW:
    *value = c->winForce;
    return VPD_OK;
  case VPD_TUNE_DEFGRAD_WINDOW_BONDS:
    *value = c->takeDefgrad;
    return VPD_OK;
  case VPD_TUNE_FORCE_WINDOW_BONDS:
    *value = c->takeForce;
    return VPD_OK;
  case VPD_TUNE_DEFGRAD_WINDOW_PLAN:
    *value = c->modeDefgrad;
    return VPD_OK;
  case VPD_TUNE_FORCE_WINDOW_PLAN:
    *value = c->modeForce;
    return VPD_OK;
  default:
    return fail(c, VPD_ERR_ARGUMENT, "Unknown tuning knob.");
  }
}

// Times both bond passes over a grid of (sweeps per pass) × (workgroups per CU) × (gather
// window), then the windows with per-window bond limits, the planned forms and the group
// windows, and keeps the fastest for each. Each pass is timed in the form the steps use, and the timing runs leave
// the state as it was: F goes to its scratch plane (with the stretch test when damage is on;
// the bond flags and marks are restored afterwards), and the force pass integrates copies of
// u, v and a.
int vpd_autotune(vpd_context *c, int reps, int *tuned) {
  if (int s = require_ready(c))
    return s;
  if (reps < 1)
    reps = 3;
  if (int s = ensure_outputs(c))
    return s;
  // LDS reservations giving at most 8 (no limit), 4, 3 and 2 workgroups per CU.
  const int ldsChoices[] = {0, 36408, 50000, 64000};
  const int roundChoices[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32};
  hipEvent_t a = nullptr, b = nullptr;
  double *su = nullptr, *sv = nullptr, *sa = nullptr;
  uint8_t *flush = nullptr, *savedAny = nullptr;
  uint32_t *savedFlags = nullptr;
  const int64_t flagBytes = sizeof(uint32_t) * std::max<int64_t>(1, c->brokenWords);
  const int64_t anyBytes = std::max<int64_t>(1, c->n);
  bool trialsStarted = false;
  // Whatever way the function leaves (an error included), the bond flags and marks the F-pass
  // trials may have set are restored, the F scratch is marked stale once a trial has run, and
  // the events and scratch buffers go.
  struct Release {
    std::function<void()> f;
    ~Release() { f(); }
  } release{[&] {
    if (savedFlags && trialsStarted) {
      (void)hipMemcpyAsync(c->broken, savedFlags, flagBytes, hipMemcpyDeviceToDevice,
                           c->stream);
      (void)hipMemcpyAsync(c->anyBroken, savedAny, anyBytes, hipMemcpyDeviceToDevice,
                           c->stream);
    }
    if (trialsStarted)
      c->haveFnew = false; // F scratch holds the tuning runs' values
    if (a)
      (void)hipEventDestroy(a);
    if (b)
      (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(c->stream);
    dfree(su);
    dfree(sv);
    dfree(sa);
    dfree(flush);
    dfree(savedFlags);
    dfree(savedAny);
  }};
  HIPTRY(c, hipEventCreate(&a));
  HIPTRY(c, hipEventCreate(&b));
  // The force pass is timed in the form the steps use (integrating, with the next prediction
  // fused), on copies of u, v and a so that the state is left as it was.
  if (dalloc(c, su, 4 * c->n) || dalloc(c, sv, 3 * c->n) || dalloc(c, sa, 3 * c->n))
    return VPD_ERR_DEVICE;
  HIPTRY(c, hipMemcpyAsync(su, c->u, sizeof(double) * 4 * c->n, hipMemcpyDeviceToDevice,
                           c->stream));
  HIPTRY(c, hipMemcpyAsync(sv, c->v, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice,
                           c->stream));
  HIPTRY(c, hipMemcpyAsync(sa, c->a, sizeof(double) * 3 * c->n, hipMemcpyDeviceToDevice,
                           c->stream));
  ForceArgs g{};
  g.n = c->n;
  g.nblocks = ceil_div(c->n, 256);
  g.x = planes(c->x, c->n);
  g.u = UPairs{su, c->n};
  g.v = planes(sv, c->n);
  g.a = planes(sa, c->n);
  g.amap = pairs(c->amap, c->n);
  g.xv = pairs(c->xv, c->n);
  g.forceOut = planes(nullptr, c->n);
  g.vol = c->vol;
  g.bonds = sell_of(c);
  g.broken = c->broken;
  g.anyBroken = c->anyBroken;
  g.codes = c->codes;
  g.dt = 1.0e-9;
  g.density = c->density;
  g.rampNow = Ramp{0, 0, 0};
  g.rampNext = Ramp{0, 0, 0};
  g.diag = c->diag;
  g.batch = c->batch;
  const bool dmg = c->damageMode == VPD_DAMAGE_CRITICAL_STRETCH;
  // The F pass is timed in the form the steps use (with the stretch test when damage is on);
  // the flags it may set are restored afterwards.
  if (dmg) {
    if (dalloc(c, savedFlags, std::max<int64_t>(1, c->brokenWords)) ||
        dalloc(c, savedAny, anyBytes))
      return VPD_ERR_DEVICE;
    HIPTRY(c, hipMemcpyAsync(savedFlags, c->broken, flagBytes, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPTRY(c, hipMemcpyAsync(savedAny, c->anyBroken, anyBytes, hipMemcpyDeviceToDevice,
                             c->stream));
  }
  // Gather windows (0: direct gathers); a windowed pass reserves at least the LDS of its
  // windows, so smaller caps repeat the uncapped setting and are skipped.
  const int windowChoices[] = {0, 64, 80, 96, 128, 160};
  const int takeChoices[] = {5, 6, 7, 8, 10, 12};
  const int pairsOf[2] = {4, 7};
  float best[2] = {1e30f, 1e30f};
  int pickLds[2] = {0, 0}, pickRounds[2] = {1, 1};
  Gather pick[2];
  struct Candidate {
    int lds, rounds;
    Gather gx;
    float ms;
  };
  std::vector<Candidate> seen[2];
  // Before every timed launch the caches are flushed by writing a buffer larger than the L2s
  // and the 256 MB MALL together, and then the planes that the kernel before the pass in a
  // step touches are read once (X, V and u before the F pass: the force pass gathers X and V
  // and writes u; A before the force pass, written by the material update): in a step the
  // other kernels evict the rest of a pass's data between its launches, and a pass timed on
  // warm caches ranks the settings differently.
  const int64_t flushBytes = int64_t(512) << 20;
  if (dalloc(c, flush, flushBytes))
    return VPD_ERR_DEVICE;
  // Mean time of `runs` launches after a warming one.
  auto time_setting = [&](int pass, int lds, int rounds, const Gather &gx, int runs,
                          float *mean, float *fastest = nullptr) -> int {
    g.rd = rounds_of(c->n, rounds);
    float total = 0.f, low = 1e30f;
    for (int r = 0; r <= runs; ++r) {
      HIPTRY(c, hipMemsetAsync(flush, r & 0xFF, flushBytes, c->stream));
      const int64_t planeBytes = (int64_t)sizeof(double) * 4 * c->n; // u, or X and V
      if (pass == 0) {
        HIPTRY(c, hipMemcpyAsync(flush, c->xv, std::min(planeBytes, flushBytes / 2),
                                 hipMemcpyDeviceToDevice, c->stream));
        HIPTRY(c, hipMemcpyAsync(flush + flushBytes / 2, c->u,
                                 std::min(planeBytes, flushBytes / 2), hipMemcpyDeviceToDevice,
                                 c->stream));
      } else {
        HIPTRY(c, hipMemcpyAsync(flush, c->amap, std::min(planeBytes * 10 / 4, flushBytes),
                                 hipMemcpyDeviceToDevice, c->stream));
      }
      HIPTRY(c, hipEventRecord(a, c->stream));
      trialsStarted = true;
      if (pass == 0)
        launch_defgrad_with(c, dmg, lds, rounds, c->batch, gx);
      else
        launch_force_with(c, g, dmg, true, true, lds, false, gx);
      HIPTRY(c, hipGetLastError());
      HIPTRY(c, hipEventRecord(b, c->stream));
      HIPTRY(c, hipEventSynchronize(b));
      float ms = 0.f;
      HIPTRY(c, hipEventElapsedTime(&ms, a, b));
      if (r > 0) { // the first run warms caches and code
        total += ms;
        low = std::min(low, ms);
      }
    }
    *mean = total / runs;
    if (fastest)
      *fastest = low;
    return VPD_OK;
  };
  auto keep_if_best = [&](int pass, int lds, int rounds, const Gather &gx, float ms) {
    if (ms < best[pass]) {
      best[pass] = ms;
      pickLds[pass] = lds;
      pickRounds[pass] = rounds;
      pick[pass] = gx;
    }
  };
  // One setting with `reps` runs, ranked by its fastest run (one slow run, e.g. a clock dip,
  // then cannot push a setting out of the final stage); keeps it if it is the fastest yet.
  auto trial = [&](int pass, int lds, int rounds, const Gather &gx) -> int {
    float mean = 0.f, ms = 0.f;
    if (int s = time_setting(pass, lds, rounds, gx, reps, &mean, &ms))
      return s;
    seen[pass].push_back(Candidate{lds, rounds, gx, ms});
    keep_if_best(pass, lds, rounds, gx, ms);
    return VPD_OK;
  };
  auto walk = [](int win, int take) {
    Gather gx;
    gx.window = win;
    gx.take = take;
    return gx;
  };
  // Stage 1: sweeps × LDS cap × window (walked, no bond limit per window); stage 2, at the
  // sweeps stage 1 chose: every window with each bond limit (walked, at stage 1's LDS cap),
  // every planned form, window and bond limit with each LDS cap, and the group windows with
  // each LDS cap (these two also at one sweep), and shared rows on small blocks; stage 3: sweeps × LDS
  // cap again for the form, window and limit chosen; stage 4: the fastest four settings (at
  // most two per form) timed again with four times the runs and ranked by their mean,
  // against timing noise between close ones.
  for (int pass = 0; pass < 2; ++pass) {
    for (int win : windowChoices)
      for (int lds : ldsChoices) {
        if (win > 0 && lds != 0 && lds <= window_lds(pairsOf[pass], win))
          continue;
        for (int rounds : roundChoices) {
          if (rounds > 1 && rounds > ceil_div(c->n, 256))
            continue;
          if (int s = trial(pass, lds, rounds, walk(win, 0)))
            return s;
        }
      }
    const int lds = pickLds[pass], rounds = pickRounds[pass];
    // The planned and group forms also run at one sweep: their staging changes which sweep
    // count suits them, and stage 3 varies the sweeps of the chosen form only.
    const std::vector<int> formRounds =
        rounds == 1 ? std::vector<int>{1} : std::vector<int>{rounds, 1};
    for (int win : windowChoices)
      for (int take : takeChoices)
        if (win > 0)
          if (int s = trial(pass, lds, rounds, walk(win, take)))
            return s;
    for (int win : kPlannedWindows)
      for (int take = 5; take <= kPlanTake; ++take) {
        if (int s = build_plan(c, pass, win, take))
          return s;
        for (int mode = 1; mode <= 2; ++mode)
          for (int cap : ldsChoices) {
            // A cap at or below the pass's own window LDS repeats the uncapped launch.
            const int own = window_lds(pairsOf[pass], win) * (mode == 2 ? 2 : 1);
            if (cap != 0 && cap <= own)
              continue;
            Gather gx = walk(win, take);
            gx.mode = mode;
            gx.plan = plan_of(c, pass);
            for (int r : formRounds)
              if (int s = trial(pass, cap, r, gx))
                return s;
          }
      }
    for (int win : kGroupWindows) {
      if (group_lds(pairsOf[pass], win) > kMaxGroupLds)
        continue;
      Gather gx = walk(win, 0);
      gx.mode = 3;
      if (int s = ensure_plan(c, pass, gx))
        return s;
      gx = with_plan(c, pass, gx);
      for (int cap : ldsChoices) {
        if (cap != 0 && cap <= group_lds(pairsOf[pass], win))
          continue;
        for (int r : formRounds)
          if (int s = trial(pass, cap, r, gx))
            return s;
      }
    }
    if (c->n <= kRowBlockMax) // shared rows: for blocks that leave most lanes idle
      for (int lanes : kRowLanes) {
        Gather gx = walk(lanes, 0);
        gx.mode = 4;
        if (int s = trial(pass, 0, 1, gx))
          return s;
      }
    Gather chosen = pick[pass];
    if (int s = ensure_plan(c, pass, chosen))
      return s;
    chosen = with_plan(c, pass, chosen);
    const int own = chosen.window == 0   ? 0
                    : chosen.mode == 3   ? group_lds(pairsOf[pass], chosen.window)
                                         : window_lds(pairsOf[pass], chosen.window) *
                                             (chosen.mode == 2 ? 2 : 1);
    for (int cap : ldsChoices)
      for (int r : roundChoices) {
        if ((cap != 0 && cap <= own) || (r > 1 && r > ceil_div(c->n, 256)) ||
            (chosen.mode == 4 && r > 1)) // shared rows launch one sweep
          continue;
        if (int s = trial(pass, cap, r, chosen))
          return s;
      }
    auto &cand = seen[pass];
    std::stable_sort(cand.begin(), cand.end(),
                     [](const Candidate &x, const Candidate &y) { return x.ms < y.ms; });
    auto same = [](const Candidate &x, const Candidate &y) {
      return x.lds == y.lds && x.rounds == y.rounds && x.gx.window == y.gx.window &&
             x.gx.take == y.gx.take && x.gx.mode == y.gx.mode;
    };
    // The fastest four distinct settings, at most two of one gather form, so that the final
    // stage (and a caller timing the candidates in steps) sees more than one form.
    std::vector<Candidate> top;
    for (const Candidate &x : cand) {
      if (top.size() == 4)
        break;
      const auto ofForm = std::count_if(top.begin(), top.end(), [&](const Candidate &y) {
        return y.gx.mode == x.gx.mode && (y.gx.window > 0) == (x.gx.window > 0);
      });
      if (ofForm < 2 &&
          std::none_of(top.begin(), top.end(), [&](const Candidate &y) { return same(x, y); }))
        top.push_back(x);
    }
    // Within 3% of the fastest the timings do not separate the settings reliably; of those,
    // the one with the fewest sweeps is kept (several sweeps ranked level in the trials have
    // run slower in steps).
    float fastest = 1e30f;
    for (Candidate &x : top) {
      if (int s = ensure_plan(c, pass, x.gx))
        return s;
      x.gx = with_plan(c, pass, x.gx);
      if (int s = time_setting(pass, x.lds, x.rounds, x.gx, 4 * reps, &x.ms))
        return s;
      fastest = std::min(fastest, x.ms);
    }
    std::stable_sort(top.begin(), top.end(),
                     [](const Candidate &x, const Candidate &y) { return x.ms < y.ms; });
    c->tunedTop[pass].clear();
    for (const Candidate &x : top)
      c->tunedTop[pass].push_back(vpd_context::TunedSetting{
          x.lds, x.rounds, x.gx.window, x.gx.take, x.gx.mode, x.ms});
    best[pass] = 1e30f;
    int fewest = INT_MAX;
    for (const Candidate &x : top)
      if (x.ms <= 1.03f * fastest && (x.rounds < fewest || (x.rounds == fewest &&
                                                            x.ms < best[pass]))) {
        fewest = x.rounds;
        best[pass] = x.ms;
        pickLds[pass] = x.lds;
        pickRounds[pass] = x.rounds;
        pick[pass] = x.gx;
      }
  }
  if (savedFlags) {
    HIPTRY(c, hipMemcpyAsync(c->broken, savedFlags, flagBytes, hipMemcpyDeviceToDevice,
                             c->stream));
    HIPTRY(c, hipMemcpyAsync(c->anyBroken, savedAny, anyBytes, hipMemcpyDeviceToDevice,
                             c->stream));
  }
  HIPTRY(c, hipStreamSynchronize(c->stream));
  trialsStarted = false; // restored here; nothing left for the release
  c->ldsDefgrad = pickLds[0];
  c->roundsDefgrad = pickRounds[0];
  c->ldsForce = pickLds[1];
  c->roundsForce = pickRounds[1];
  c->winDefgrad = pick[0].window;
  c->winForce = pick[1].window;
  c->takeDefgrad = pick[0].take;
  c->takeForce = pick[1].take;
  c->modeDefgrad = pick[0].mode;
  c->modeForce = pick[1].mode;
  c->haveFnew = false; // F scratch now holds the tuning runs' values
  if (int s = prepare_plans(c)) // the plans of the picks (the trials replaced them)
    return s;
  if (tuned) {
    tuned[VPD_TUNE_DEFGRAD_LDS] = c->ldsDefgrad;
    tuned[VPD_TUNE_FORCE_LDS] = c->ldsForce;
    tuned[VPD_TUNE_DEFGRAD_ROUNDS] = c->roundsDefgrad;
    tuned[VPD_TUNE_FORCE_ROUNDS] = c->roundsForce;
    tuned[VPD_TUNE_DEFGRAD_WINDOW] = c->winDefgrad;
    tuned[VPD_TUNE_FORCE_WINDOW] = c->winForce;
    tuned[VPD_TUNE_DEFGRAD_WINDOW_BONDS] = c->takeDefgrad;
    tuned[VPD_TUNE_FORCE_WINDOW_BONDS] = c->takeForce;
    tuned[VPD_TUNE_DEFGRAD_WINDOW_PLAN] = c->modeDefgrad;
    tuned[VPD_TUNE_FORCE_WINDOW_PLAN] = c->modeForce;
  }
  return check_device(c);
}

int vpd_set_return_map(vpd_context *c, int mode, double rtol) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  if (mode < RM_BISECTION || mode > RM_LEARNED)
    return fail(c, VPD_ERR_ARGUMENT, "Unknown return-map mode.");
  if (!(rtol > 0.0) || rtol > 1e-3)
    return fail(c, VPD_ERR_ARGUMENT, "Return-map tolerance must lie in (0, 1e-3].");
  c->returnMap = mode;
  c->rtol = rtol;
  return VPD_OK;
}

int vpd_attach_warm_start(vpd_context *c, vpd_mlp *mlp) {
  if (!c)
    return fail(c, VPD_ERR_ARGUMENT, "Null vpd_context.");
  c->mlp = mlp;
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
  if (enable < 0)
    return fail(c, VPD_ERR_ARGUMENT, "The timing stride must be 0 (off) or positive.");
  c->timing = enable != 0;
  c->timingStride = std::max(enable, 1);
  c->timingTick = 0;
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

} // extern "C"

// ---- allocation-free neighbour build ---------------------------------------------------
// The work memory of vpd_build_neighbors_noalloc is one file-scope device array, `scratch`;
// every GPU the library runs on holds its own copy, at its own address, until the process
// ends. In host code the array's name is not a device address: each device's address comes
// from hipGetSymbolAddress (scratch_base) and only that address is handed to kernels.
// Layout, for at most kNoAllocMaxParticles particles and kNoAllocMaxCells cells:
//   Bounds                                       (256 B)
//   part      int64[max(n + 1, cells)]            scan partial sums
//   totals    int64[ceil(max(n + 1, cells) / SCAN_ITEMS) + 1]   scan block totals
//   cellOf    int32[n + 1]   cell of each particle; reused for the row counts
//   items     int32[n]       particle ids by cell
//   cellCount, cellStart, cursor   int32[cells] each
namespace {

constexpr int64_t kNoAllocMaxParticles = int64_t(1) << 21; // 2,097,152
constexpr int64_t kNoAllocMaxCells = int64_t(1) << 20;     // 1,048,576
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kScanItems = (size_t)kNoAllocMaxParticles + 1; // >= kNoAllocMaxCells
constexpr size_t kOffPart = 256;
constexpr size_t kOffTotals = kOffPart + align256(8 * kScanItems);
constexpr size_t kOffCellOf =
    kOffTotals + align256(8 * ((kScanItems + SCAN_ITEMS - 1) / SCAN_ITEMS + 1));
constexpr size_t kOffItems = kOffCellOf + align256(4 * ((size_t)kNoAllocMaxParticles + 1));
constexpr size_t kOffCells = kOffItems + align256(4 * (size_t)kNoAllocMaxParticles);
constexpr size_t kScratchBytes = kOffCells + 3 * align256(4 * (size_t)kNoAllocMaxCells);
static_assert(sizeof(Bounds) <= 256, "Bounds must fit its slot");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchBase[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with that device
// current (hipGetSymbolAddress; the array's name itself is not a device address here).
int scratch_base(unsigned char **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(nullptr, VPD_ERR_UNSUPPORTED, "Device index out of range.");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (!g_scratchBase[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, VPD_ERR_DEVICE,
                  std::string("HIP error: ") + hipGetErrorString(e) + " at hipGetSymbolAddress");
    g_scratchBase[dev] = p;
  }
  *out = static_cast<unsigned char *>(g_scratchBase[dev]);
  return VPD_OK;
}

__global__ void __launch_bounds__(256) k_fill_int(int *p, int64_t count, int value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count)
    p[i] = value;
}

} // namespace

extern "C" {

int vpd_build_neighbors_noalloc(const double *x, int64_t n, double horizon, int32_t *rowptr,
                                int32_t *col, int64_t colCapacity, int64_t *nnzOut,
                                void *stream) {
  if (!x || !rowptr || !nnzOut || n < 0 || colCapacity < 0 || (colCapacity > 0 && !col))
    return fail(nullptr, VPD_ERR_ARGUMENT, "Invalid neighbour-build arguments.");
  *nnzOut = 0;
  if (n > kNoAllocMaxParticles)
    return fail(nullptr, VPD_ERR_ARGUMENT,
                "The allocation-free neighbour build takes at most 2097152 particles.");
  if (!(horizon > 0.0) || !std::isfinite(horizon))
    return fail(nullptr, VPD_ERR_ARGUMENT, "PD horizon must be positive and finite.");
  hipStream_t st = static_cast<hipStream_t>(stream);
  auto hipfail = [](hipError_t e) {
    return fail(nullptr, VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
  };
#define NATRY(expr)                                                                            \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return hipfail(e_);                                                                      \
  } while (0)
  if (n == 0) {
    const int32_t zero = 0;
    NATRY(hipMemcpyAsync(rowptr, &zero, sizeof(int32_t), hipMemcpyHostToDevice, st));
    NATRY(hipStreamSynchronize(st));
    return VPD_OK;
  }
  unsigned char *base = nullptr;
  if (int s = scratch_base(&base))
    return s;
  Bounds *db = reinterpret_cast<Bounds *>(base);
  int64_t *part = reinterpret_cast<int64_t *>(base + kOffPart);
  int64_t *totals = reinterpret_cast<int64_t *>(base + kOffTotals);
  int *cellOf = reinterpret_cast<int *>(base + kOffCellOf);
  int *items = reinterpret_cast<int *>(base + kOffItems);
  int *cellCount = reinterpret_cast<int *>(base + kOffCells);
  int *cellStart = reinterpret_cast<int *>(base + kOffCells + align256(4 * (size_t)kNoAllocMaxCells));
  int *cursor = reinterpret_cast<int *>(base + kOffCells + 2 * align256(4 * (size_t)kNoAllocMaxCells));

  // Search radius and cell grid as in vpd_build_neighbors (PdCorrespondenceSkeleton.cpp:30-39,
  // 101-102).
  Bounds hb;
  hb.maxAbsBits = 0;
  for (int d = 0; d < 3; ++d) {
    hb.cmin[d] = LLONG_MAX;
    hb.cmax[d] = LLONG_MIN;
  }
  hb.nonfinite = 0;
  NATRY(hipMemcpyAsync(db, &hb, sizeof(Bounds), hipMemcpyHostToDevice, st));
  k_abs_max<<<grid_of(n), 256, 0, st>>>(n, x, db);
  NATRY(hipGetLastError());
  NATRY(hipMemcpyAsync(&hb, db, sizeof(Bounds), hipMemcpyDeviceToHost, st));
  NATRY(hipStreamSynchronize(st));
  if (hb.nonfinite)
    return fail(nullptr, VPD_ERR_NUMERIC, "Nonfinite PD coordinate.");
  double maxAbs;
  std::memcpy(&maxAbs, &hb.maxAbsBits, sizeof(double));
  const double scale = std::max(horizon, maxAbs);
  const double tol = 64.0 * std::numeric_limits<double>::epsilon() * scale;
  Grid g;
  g.radius = horizon + tol;
  g.cw = horizon + 2.0 * tol;
  k_cell_bounds<<<grid_of(n), 256, 0, st>>>(n, x, g.cw, db);
  NATRY(hipGetLastError());
  NATRY(hipMemcpyAsync(&hb, db, sizeof(Bounds), hipMemcpyDeviceToHost, st));
  NATRY(hipStreamSynchronize(st));
  double cells = 1.0;
  for (int d = 0; d < 3; ++d) {
    g.c0[d] = hb.cmin[d];
    g.dim[d] = hb.cmax[d] - hb.cmin[d] + 1;
    cells *= (double)g.dim[d];
  }
  if (cells > (double)kNoAllocMaxCells)
    return fail(nullptr, VPD_ERR_UNSUPPORTED,
                "Cell grid too large for the allocation-free neighbour build (more than "
                "1048576 cells).");
  const int64_t nc = (int64_t)cells;

  // Cells, then per-particle counts -> rowptr, then rows; every counter starts at zero.
  k_fill_int<<<grid_of(nc), 256, 0, st>>>(cellCount, nc, 0);
  k_fill_int<<<grid_of(nc), 256, 0, st>>>(cursor, nc, 0);
  k_cell_count<<<grid_of(n), 256, 0, st>>>(n, x, g, cellOf, cellCount);
  NATRY(hipGetLastError());
  int64_t placed = 0;
  NATRY(scan_ints_in(st, cellCount, cellStart, nc, part, totals, &placed));
  k_cell_fill<<<grid_of(n), 256, 0, st>>>(n, cellOf, cellStart, cursor, items);
  int *rowCount = cellOf; // cellOf is not read again
  k_fill_int<<<grid_of(n + 1), 256, 0, st>>>(rowCount, n + 1, 0);
  k_row_scan<false><<<grid_of(n), 256, 0, st>>>(n, x, g, cellStart, cellCount, items, rowCount,
                                                nullptr, nullptr);
  NATRY(hipGetLastError());
  int64_t nnz = 0;
  NATRY(scan_ints_in(st, rowCount, rowptr, n + 1, part, totals, &nnz));
  if (nnz > INT_MAX)
    return fail(nullptr, VPD_ERR_UNSUPPORTED, "Neighbour count exceeds the int32 CSR range.");
  *nnzOut = nnz;
  if (nnz > colCapacity)
    return fail(nullptr, VPD_ERR_ARGUMENT,
                "Neighbour list capacity too small: " + std::to_string(nnz) + " entries needed.");
  k_row_scan<true><<<grid_of(n), 256, 0, st>>>(n, x, g, cellStart, cellCount, items, nullptr,
                                               rowptr, col);
  k_row_sort<<<(int)n, 64, 0, st>>>(n, rowptr, col);
  NATRY(hipGetLastError());
  NATRY(hipStreamSynchronize(st));
#undef NATRY
  return VPD_OK;
}

// The largest errors of the two FP32 hardware steps of the bisection's yield screen
// (jc_yield_screen) over every input the screen can give them: v_log_f32 on every float in
// [0.5, 1] (absolute, against the FP64 log2; 1 is where a mantissa just below 1 rounds) and
// v_exp_f32 on every float in [0, 1] (relative, against the FP64 exp2). Each workgroup reduces its share and makes one atomic max (the bits
// of non-negative doubles order like their values).
__device__ __forceinline__ void worst_of_group(double e, unsigned long long *worst) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1)
    e = fmax(e, __shfl_xor(e, off));
  if ((threadIdx.x & 63) == 0)
    red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(worst, (unsigned long long)__double_as_longlong(
                         fmax(fmax(red[0], red[1]), fmax(red[2], red[3]))));
}
__global__ void __launch_bounds__(256) k_screen_log2_error(unsigned long long *worst) {
  double e = 0.0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i <= (1u << 23); i += gridDim.x * 256) {
    const float m = __uint_as_float(0x3F000000u + i); // 0.5·(1 + i·2⁻²³)
    e = fmax(e, fabs((double)__builtin_amdgcn_logf(m) - log2((double)m)));
  }
  worst_of_group(e, worst);
}
__global__ void __launch_bounds__(256) k_screen_exp2_error(unsigned long long *worst) {
  double e = 0.0;
  for (uint32_t b = blockIdx.x * 256 + threadIdx.x; b <= 0x3F800000u; b += gridDim.x * 256) {
    const float f = __uint_as_float(b); // every float in [0, 1]
    const double ref = exp2((double)f);
    e = fmax(e, fabs((double)__builtin_amdgcn_exp2f(f) - ref) / ref);
  }
  worst_of_group(e, worst);
}

int vpd_screen_accuracy(double *log2_abs_error, double *exp2_rel_error) {
  if (!log2_abs_error || !exp2_rel_error)
    return fail(nullptr, VPD_ERR_ARGUMENT, "Invalid accuracy arguments.");
  unsigned long long *worst = nullptr;
  if (hipMalloc((void **)&worst, 2 * sizeof(unsigned long long)) != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, "HIP allocation failed.");
  unsigned long long host[2] = {0ull, 0ull};
  hipError_t e = hipMemcpy(worst, host, sizeof(host), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_screen_log2_error<<<1024, 256>>>(worst);
    k_screen_exp2_error<<<4096, 256>>>(worst + 1);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(host, worst, sizeof(host), hipMemcpyDeviceToHost);
  (void)hipFree(worst);
  if (e != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
  std::memcpy(log2_abs_error, &host[0], sizeof(double));
  std::memcpy(exp2_rel_error, &host[1], sizeof(double));
  return VPD_OK;
}

static int return_map_run(int mode, int64_t nblock, int nstatev, const double *props, int nprops,
                          double dt, double rtol, const double *strainInc,
                          const double *stressOld, const double *stateOld,
                          const double *enerInternOld, const double *enerInelasOld,
                          const double *seed, double *stressNew, double *stateNew,
                          double *enerInternNew, double *enerInelasNew, int32_t *iterations,
                          int32_t *fallbacks, double *elapsedMs, vpd_mlp *net = nullptr,
                          double *seedOut = nullptr) {
  if (mode < RM_BISECTION || mode > VPD_RETURN_MAP_BISECTION_GROUPED)
    return fail(nullptr, VPD_ERR_ARGUMENT, "Unknown return-map mode.");
  if (!props || nprops < 14 || nstatev < 6)
    return fail(nullptr, VPD_ERR_ARGUMENT, "JohnsonCook VUMAT requires 14 props and 6 states");
  if (mode == RM_LEARNED && !seed && !net)
    return fail(nullptr, VPD_ERR_ARGUMENT, "The learned return map needs a seed array.");
  if (!(dt > 0.0))
    return fail(nullptr, VPD_ERR_ARGUMENT, "VUMAT update requires dt > 0.");
  if (nblock <= 0)
    return VPD_OK;
  VumatArgs g{};
  g.nblock = nblock;
  g.nstatev = nstatev;
  g.nfieldv = 0;
  g.jc = make_jc(std::vector<double>(props, props + nprops));
  g.dt = dt;
  g.rtol = rtol;
  g.strainInc = strainInc;
  g.stressOld = stressOld;
  g.stateOld = stateOld;
  g.enerInternOld = enerInternOld;
  g.enerInelasOld = enerInelasOld;
  g.seed = seed;
  double *seeds = nullptr; // the learned mode with the network: its seeds, then the solve
  if (net) {
    if (!seedOut && nblock > net->seedCap) {
      dfree(net->seedBuf);
      net->seedCap = 0;
      if (hipMalloc((void **)&net->seedBuf, sizeof(double) * (size_t)nblock) != hipSuccess)
        return fail(nullptr, VPD_ERR_DEVICE, "HIP allocation failed.");
      net->seedCap = nblock;
    }
    seeds = seedOut ? seedOut : net->seedBuf;
    g.seed = seeds;
    g.seedFromState5 = seedOut == nullptr; // without seed_out only the network's are written
  }
  g.stressNew = stressNew;
  g.stateNew = stateNew;
  g.enerInternNew = enerInternNew;
  g.enerInelasNew = enerInelasNew;
  g.iterations = iterations;
  g.fallbacks = fallbacks;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, "HIP event creation failed.");
  (void)hipEventRecord(a, nullptr);
  switch (mode) {
  case RM_BISECTION:
    k_vumat_jc<RM_BISECTION><<<grid_of(nblock), 256>>>(g);
    break;
  case RM_BISECTION_TIGHT:
    k_vumat_jc<RM_BISECTION_TIGHT><<<grid_of(nblock), 256>>>(g);
    break;
  case RM_NEWTON:
    k_vumat_jc<RM_NEWTON><<<grid_of(nblock), 256>>>(g);
    break;
  case VPD_RETURN_MAP_BISECTION_UNSCREENED:
    k_vumat_jc<RM_BISECTION_UNSCREENED><<<grid_of(nblock), 256>>>(g);
    break;
  case VPD_RETURN_MAP_BISECTION_GROUPED:
    k_vumat_jc_g<kListLanes><<<(int)ceil_div(nblock * kListLanes, 256), 256>>>(g);
    break;
  default:
    if (net)
      k_seed_block<<<(int)ceil_div(nblock, 2048), 256>>>(nblock, strainInc, nblock, stateOld,
                                                          nblock, net->theta, net->norm, seeds,
                                                          seedOut != nullptr);
    k_vumat_jc<RM_LEARNED><<<grid_of(nblock), 256>>>(g);
  }
  hipError_t e = hipGetLastError();
  (void)hipEventRecord(b, nullptr);
  if (e == hipSuccess)
    e = hipEventSynchronize(b);
  float ms = 0.f;
  if (e == hipSuccess)
    e = hipEventElapsedTime(&ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (e != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
  if (elapsedMs)
    *elapsedMs = ms;
  return VPD_OK;
}

int vpd_return_map_device(int mode, int64_t nblock, int nstatev, const double *props, int nprops,
                          double dt, double rtol, const double *strainInc,
                          const double *stressOld, const double *stateOld,
                          const double *enerInternOld, const double *enerInelasOld,
                          const double *seed, double *stressNew, double *stateNew,
                          double *enerInternNew, double *enerInelasNew, int32_t *iterations,
                          int32_t *fallbacks, double *elapsedMs) {
  return return_map_run(mode, nblock, nstatev, props, nprops, dt, rtol, strainInc, stressOld,
                        stateOld, enerInternOld, enerInelasOld, seed, stressNew, stateNew,
                        enerInternNew, enerInelasNew, iterations, fallbacks, elapsedMs);
}

int vpd_return_map_learned_device(vpd_mlp *mlp, int64_t nblock, int nstatev,
                                  const double *props, int nprops, double dt, double rtol,
                                  const double *strainInc, const double *stressOld,
                                  const double *stateOld, const double *enerInternOld,
                                  const double *enerInelasOld, double *seedOut,
                                  double *stressNew, double *stateNew, double *enerInternNew,
                                  double *enerInelasNew, int32_t *iterations, int32_t *fallbacks,
                                  double *elapsedMs) {
  if (!mlp)
    return fail(nullptr, VPD_ERR_ARGUMENT, "The learned return map needs a warm-start network.");
  return return_map_run(RM_LEARNED, nblock, nstatev, props, nprops, dt, rtol, strainInc,
                        stressOld, stateOld, enerInternOld, enerInelasOld, nullptr, stressNew,
                        stateNew, enerInternNew, enerInelasNew, iterations, fallbacks, elapsedMs,
                        mlp, seedOut);
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
  // VumatMaterialAdapter::updateBlock (src/material/VumatMaterialAdapter.cpp:150-151).
  if (!std::isfinite(*dt) || *dt <= 0.0)
    return fail(nullptr, VPD_ERR_ARGUMENT, "VUMAT update requires dt > 0.");
  if (*nfieldv < 0)
    return fail(nullptr, VPD_ERR_ARGUMENT, "nfieldv must be non-negative.");
  if (*nblock <= 0)
    return VPD_OK;
  // Arrays the routine reads or writes (JohnsonCook.f90:74-155; the *New copies of stretch,
  // deformation gradient and field variables read their *Old arrays).
  if (!strainInc || !stressOld || !stateOld || !enerInternOld || !enerInelasOld || !stressNew ||
      !stateNew || !enerInternNew || !enerInelasNew || (stretchNew && !stretchOld) ||
      (defgradNew && !defgradOld) || (fieldNew && *nfieldv > 0 && !fieldOld))
    return fail(nullptr, VPD_ERR_ARGUMENT, "Null VUMAT array.");
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
  g.rtol = 0.0;
  g.seed = nullptr;
  g.seedFromState5 = false;
  g.iterations = nullptr;
  g.fallbacks = nullptr;
  k_vumat_jc<RM_BISECTION><<<grid_of(g.nblock), 256>>>(g);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return fail(nullptr, VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
  return VPD_OK;
}

} // extern "C"

// ---------------------------------------------------------------- warm-start network


namespace {

int mfail(vpd_mlp *m, int status, const std::string &msg) {
  if (m)
    m->err = msg;
  g_error = msg;
  return status;
}

#define MLPTRY(m, expr)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return mfail((m), VPD_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_));   \
  } while (0)

template <class T> int malloc_n(vpd_mlp *m, T *&p, int64_t count) {
  if (p)
    (void)hipFree(p);
  p = nullptr;
  MLPTRY(m, hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * (size_t)std::max<int64_t>(1, count)));
  return VPD_OK;
}

// Forward and backward over the batch in xb/yb; leaves the gradient in m->grad and the
// per-sample squared errors in m->sq.
void mlp_forward_backward(vpd_mlp *m, int B) {
  hipStream_t s = m->stream;
  float *t = m->theta, *g = m->grad;
  const int H = MLP_H;
  auto grid = [](int64_t n) { return (int)ceil_div(n, 256); };
  k_mlp_dense<<<grid((int64_t)B * H), 256, 0, s>>>(B, MLP_IN, H, m->xb, t + MLP_OFF_W1,
                                                    t + MLP_OFF_B1, m->h1, 1);
  k_mlp_dense<<<grid((int64_t)B * H), 256, 0, s>>>(B, H, H, m->h1, t + MLP_OFF_W2, t + MLP_OFF_B2,
                                                    m->h2, 1);
  k_mlp_dense<<<grid((int64_t)B * H), 256, 0, s>>>(B, H, H, m->h2, t + MLP_OFF_W3, t + MLP_OFF_B3,
                                                    m->h3, 1);
  k_mlp_dense<<<grid(B), 256, 0, s>>>(B, H, 1, m->h3, t + MLP_OFF_W4, t + MLP_OFF_B4, m->y, 0);
  k_mlp_loss_grad<<<grid(B), 256, 0, s>>>(B, m->y, m->yb, m->dy, m->sq);
  const int chunks = (int)ceil_div(B, MLP_CHUNK);
  auto wgrad = [&](int K, int N, const float *dZ, const float *X, int offW, int offB) {
    dim3 gp((unsigned)ceil_div((int64_t)N * (K + 1), 256), (unsigned)chunks);
    k_mlp_wgrad_part<<<gp, 256, 0, s>>>(B, K, N, dZ, X, m->part);
    k_mlp_wgrad_reduce<<<grid((int64_t)N * (K + 1)), 256, 0, s>>>(chunks, K, N, m->part, g + offW,
                                                                 g + offB);
  };
  wgrad(H, 1, m->dy, m->h3, MLP_OFF_W4, MLP_OFF_B4);
  k_mlp_back_input<<<grid((int64_t)B * H), 256, 0, s>>>(B, H, 1, m->dy, t + MLP_OFF_W4, m->h3,
                                                         m->dz3);
  wgrad(H, H, m->dz3, m->h2, MLP_OFF_W3, MLP_OFF_B3);
  k_mlp_back_input<<<grid((int64_t)B * H), 256, 0, s>>>(B, H, H, m->dz3, t + MLP_OFF_W3, m->h2,
                                                         m->dz2);
  wgrad(H, H, m->dz2, m->h1, MLP_OFF_W2, MLP_OFF_B2);
  k_mlp_back_input<<<grid((int64_t)B * H), 256, 0, s>>>(B, H, H, m->dz2, t + MLP_OFF_W2, m->h1,
                                                         m->dz1);
  wgrad(MLP_IN, H, m->dz1, m->xb, MLP_OFF_W1, MLP_OFF_B1);
}

// The bf16 MFMA step over the batch in xb/yb (see vpd_mlp.hpp): the gradient to m->grad and,
// with `adam`, the Adam update; per-point squared errors to m->sq.
void mlp_step_mfma(vpd_mlp *m, int B, bool adam, float lr, float c1, float c2) {
  hipStream_t s = m->stream;
  const int64_t ld = ceil_div(B, MLP_GRAD_CHUNK) * MLP_GRAD_CHUNK;
  MlpTrainBufs b{B, ld, m->xb, m->yb, m->xT, m->h1T, m->h2T, m->h3T, m->dz1T, m->dz2T, m->dz3T,
                 m->dyT, m->sq};
  const int64_t tiles = ld / 16, chunks = ld / MLP_GRAD_CHUNK;
  k_mlp_fwd_bwd<<<(int)std::min<int64_t>(ceil_div(tiles, 4), 256), 256, 0, s>>>(b, m->theta);
  k_mlp_wgrad<<<dim3((unsigned)ceil_div(chunks, 4), 3), 256, 0, s>>>(b, m->gpart);
  k_mlp_reduce_adam<<<(int)ceil_div(MLP_PARAMS, 256), 256, 0, s>>>(
      (int)chunks, m->gpart, m->grad, m->theta, m->m, m->v, adam ? 1 : 0, lr, c1, c2);
}

int mlp_batch_loss(vpd_mlp *m, int B, float *loss) {
  std::vector<float> sq(B);
  MLPTRY(m, hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * B, hipMemcpyDeviceToHost, m->stream));
  MLPTRY(m, hipStreamSynchronize(m->stream));
  double acc = 0.0;
  for (float v : sq)
    acc += v;
  if (loss)
    *loss = (float)(acc / B);
  return VPD_OK;
}

int mlp_alloc_mfma(vpd_mlp *m) {
  m->ld = ceil_div(m->maxB, MLP_GRAD_CHUNK) * MLP_GRAD_CHUNK;
  const int64_t ld = m->ld, H = MLP_H;
  return malloc_n(m, m->xT, MLP_IN * ld) || malloc_n(m, m->h1T, H * ld) ||
         malloc_n(m, m->h2T, H * ld) || malloc_n(m, m->h3T, H * ld) ||
         malloc_n(m, m->dz1T, H * ld) || malloc_n(m, m->dz2T, H * ld) ||
         malloc_n(m, m->dz3T, H * ld) || malloc_n(m, m->dyT, ld) ||
         malloc_n(m, m->gpart, ld / MLP_GRAD_CHUNK * MLP_PARAMS);
}

} // namespace

extern "C" {

int vpd_mlp_create(int maxBatch, vpd_mlp **out) {
  if (!out || maxBatch <= 0)
    return mfail(nullptr, VPD_ERR_ARGUMENT, "vpd_mlp_create requires a batch size and output.");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return mfail(nullptr, VPD_ERR_DEVICE, "No HIP device available.");
  auto *m = new vpd_mlp();
  m->maxB = maxBatch;
  const int64_t B = maxBatch, H = MLP_H;
  const int64_t chunks = ceil_div(B, MLP_CHUNK);
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      malloc_n(m, m->theta, MLP_PARAMS) || malloc_n(m, m->norm, MLP_NORM) ||
      malloc_n(m, m->m, MLP_PARAMS) || malloc_n(m, m->v, MLP_PARAMS) ||
      malloc_n(m, m->grad, MLP_PARAMS) || malloc_n(m, m->xb, B * MLP_IN) ||
      malloc_n(m, m->yb, B) || malloc_n(m, m->h1, B * H) || malloc_n(m, m->h2, B * H) ||
      malloc_n(m, m->h3, B * H) || malloc_n(m, m->y, B) || malloc_n(m, m->dy, B) ||
      malloc_n(m, m->dz3, B * H) || malloc_n(m, m->dz2, B * H) || malloc_n(m, m->dz1, B * H) ||
      malloc_n(m, m->part, chunks * H * (H + 1)) || malloc_n(m, m->sq, B) ||
      mlp_alloc_mfma(m)) {
    vpd_mlp_destroy(m);
    return mfail(nullptr, VPD_ERR_DEVICE, "Allocating the warm-start network failed.");
  }
  (void)hipMemsetAsync(m->theta, 0, sizeof(float) * MLP_PARAMS, m->stream);
  (void)hipMemsetAsync(m->m, 0, sizeof(float) * MLP_PARAMS, m->stream);
  (void)hipMemsetAsync(m->v, 0, sizeof(float) * MLP_PARAMS, m->stream);
  std::vector<float> norm(MLP_NORM, 0.f);
  for (int i = MLP_IN; i < 2 * MLP_IN; ++i)
    norm[i] = 1.f;
  norm[2 * MLP_IN + 1] = 1.f;
  (void)hipMemcpyAsync(m->norm, norm.data(), sizeof(float) * MLP_NORM, hipMemcpyHostToDevice,
                       m->stream);
  if (hipStreamSynchronize(m->stream) != hipSuccess) {
    vpd_mlp_destroy(m);
    return mfail(nullptr, VPD_ERR_DEVICE, "Initialising the warm-start network failed.");
  }
  *out = m;
  return VPD_OK;
}

void vpd_mlp_destroy(vpd_mlp *m) {
  if (!m)
    return;
  if (m->stream)
    (void)hipStreamSynchronize(m->stream);
  float **bufs[] = {&m->theta, &m->norm, &m->m,   &m->v,   &m->grad, &m->xb, &m->yb,
                    &m->h1,    &m->h2,   &m->h3,  &m->y,   &m->dy,   &m->dz3, &m->dz2,
                    &m->dz1,   &m->part, &m->sq,  &m->xs,  &m->ys};
  for (float **b : bufs)
    dfree(*b);
  short **sbufs[] = {&m->xT, &m->h1T, &m->h2T, &m->h3T, &m->dz1T, &m->dz2T, &m->dz3T};
  for (short **b : sbufs)
    dfree(*b);
  dfree(m->dyT);
  dfree(m->gpart);
  dfree(m->order);
  dfree(m->seedBuf);
  if (m->stream)
    (void)hipStreamDestroy(m->stream);
  delete m;
}

const char *vpd_mlp_last_error(const vpd_mlp *m) { return m ? m->err.c_str() : g_error.c_str(); }

int vpd_mlp_set_params(vpd_mlp *m, const float *theta, const float *norm) {
  if (!m)
    return mfail(m, VPD_ERR_ARGUMENT, "Null network.");
  if (theta)
    MLPTRY(m, hipMemcpyAsync(m->theta, theta, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice,
                             m->stream));
  if (norm) {
    for (int i = MLP_IN; i < 2 * MLP_IN; ++i)
      if (!(norm[i] > 0.f))
        return mfail(m, VPD_ERR_ARGUMENT, "Input scales must be positive.");
    MLPTRY(m, hipMemcpyAsync(m->norm, norm, sizeof(float) * MLP_NORM, hipMemcpyHostToDevice,
                             m->stream));
  }
  MLPTRY(m, hipMemsetAsync(m->m, 0, sizeof(float) * MLP_PARAMS, m->stream));
  MLPTRY(m, hipMemsetAsync(m->v, 0, sizeof(float) * MLP_PARAMS, m->stream));
  MLPTRY(m, hipStreamSynchronize(m->stream));
  m->adamStep = 0;
  return VPD_OK;
}

int vpd_mlp_get_params(vpd_mlp *m, float *theta, float *norm) {
  if (!m)
    return mfail(m, VPD_ERR_ARGUMENT, "Null network.");
  if (theta)
    MLPTRY(m, hipMemcpyAsync(theta, m->theta, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost,
                             m->stream));
  if (norm)
    MLPTRY(m, hipMemcpyAsync(norm, m->norm, sizeof(float) * MLP_NORM, hipMemcpyDeviceToHost,
                             m->stream));
  MLPTRY(m, hipStreamSynchronize(m->stream));
  return VPD_OK;
}

int vpd_mlp_loss_and_gradient(vpd_mlp *m, const float *x, const float *t, int B, float *loss,
                              float *grad) {
  if (!m || !x || !t || B <= 0 || B > m->maxB)
    return mfail(m, VPD_ERR_ARGUMENT, "Batch exceeds the network's capacity.");
  MLPTRY(m, hipMemcpyAsync(m->xb, x, sizeof(float) * B * MLP_IN, hipMemcpyHostToDevice,
                           m->stream));
  MLPTRY(m, hipMemcpyAsync(m->yb, t, sizeof(float) * B, hipMemcpyHostToDevice, m->stream));
  if (m->form == VPD_MLP_TRAIN_MFMA)
    mlp_step_mfma(m, B, false, 0.f, 1.f, 1.f);
  else
    mlp_forward_backward(m, B);
  MLPTRY(m, hipGetLastError());
  if (grad)
    MLPTRY(m, hipMemcpyAsync(grad, m->grad, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost,
                             m->stream));
  return mlp_batch_loss(m, B, loss);
}

int vpd_mlp_set_dataset(vpd_mlp *m, const float *x, const float *t, int64_t n, uint64_t seed) {
  if (!m || !x || !t || n <= 0 || n > INT_MAX)
    return mfail(m, VPD_ERR_ARGUMENT, "Invalid training set.");
  if (malloc_n(m, m->xs, n * MLP_IN) || malloc_n(m, m->ys, n) || malloc_n(m, m->order, n))
    return VPD_ERR_DEVICE;
  // A fixed shuffle (64-bit LCG driven Fisher-Yates) makes training reproducible.
  std::vector<int> order(n);
  for (int64_t i = 0; i < n; ++i)
    order[i] = (int)i;
  uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  for (int64_t i = n - 1; i > 0; --i) {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    const int64_t j = (int64_t)((st >> 33) % (uint64_t)(i + 1));
    std::swap(order[i], order[j]);
  }
  MLPTRY(m, hipMemcpyAsync(m->xs, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, m->stream));
  MLPTRY(m, hipMemcpyAsync(m->ys, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  MLPTRY(m, hipMemcpyAsync(m->order, order.data(), sizeof(int) * n, hipMemcpyHostToDevice,
                           m->stream));
  MLPTRY(m, hipStreamSynchronize(m->stream));
  m->nData = n;
  m->cursor = 0;
  return VPD_OK;
}

int vpd_mlp_train(vpd_mlp *m, int steps, int batch, float lr, float *lastLoss) {
  if (!m || batch <= 0 || batch > m->maxB || steps < 0)
    return mfail(m, VPD_ERR_ARGUMENT, "Batch exceeds the network's capacity.");
  if (m->nData <= 0)
    return mfail(m, VPD_ERR_STATE, "No training set.");
  for (int s = 0; s < steps; ++s) {
    k_mlp_gather<<<(int)ceil_div(batch, 256), 256, 0, m->stream>>>(batch, m->nData, m->cursor,
                                                                    m->order, m->xs, m->ys, m->xb,
                                                                    m->yb);
    m->cursor = (m->cursor + batch) % m->nData;
    ++m->adamStep;
    const float c1 = 1.f - powf(0.9f, (float)m->adamStep);
    const float c2 = 1.f - powf(0.999f, (float)m->adamStep);
    if (m->form == VPD_MLP_TRAIN_MFMA) {
      mlp_step_mfma(m, batch, true, lr, c1, c2);
      continue;
    }
    mlp_forward_backward(m, batch);
    k_mlp_adam<<<(int)ceil_div(MLP_PARAMS, 256), 256, 0, m->stream>>>(MLP_PARAMS, m->theta,
                                                                       m->grad, m->m, m->v, lr, c1,
                                                                       c2);
  }
  MLPTRY(m, hipGetLastError());
  return mlp_batch_loss(m, batch, lastLoss);
}

int vpd_mlp_set_training_form(vpd_mlp *m, int form) {
  if (!m)
    return mfail(m, VPD_ERR_ARGUMENT, "Null network.");
  if (form != VPD_MLP_TRAIN_FP32 && form != VPD_MLP_TRAIN_MFMA)
    return mfail(m, VPD_ERR_ARGUMENT, "Unknown training form.");
  m->form = form;
  return VPD_OK;
}

int vpd_mlp_predict(vpd_mlp *m, const float *x, int64_t n, float *y) {
  if (!m || !x || !y || n < 0)
    return mfail(m, VPD_ERR_ARGUMENT, "Invalid prediction arguments.");
  if (n == 0)
    return VPD_OK;
  float *dx = nullptr, *dy = nullptr;
  if (malloc_n(m, dx, n * MLP_IN) || malloc_n(m, dy, n))
    return VPD_ERR_DEVICE;
  MLPTRY(m, hipMemcpyAsync(dx, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice, m->stream));
  k_mlp_infer_raw<<<mlp_grid(n), 256, 0, m->stream>>>(n, dx, m->theta, dy);
  MLPTRY(m, hipGetLastError());
  MLPTRY(m, hipMemcpyAsync(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  MLPTRY(m, hipStreamSynchronize(m->stream));
  dfree(dx);
  dfree(dy);
  return VPD_OK;
}

int vpd_mlp_seed_device(vpd_mlp *m, const double *inc, int64_t incStride, const double *state,
                        int64_t stateStride, int64_t n, double *seed, double *elapsedMs) {
  if (!m || !inc || !state || !seed || n < 0)
    return mfail(m, VPD_ERR_ARGUMENT, "Invalid seed arguments.");
  if (n == 0)
    return VPD_OK;
  hipEvent_t a, b;
  MLPTRY(m, hipEventCreate(&a));
  MLPTRY(m, hipEventCreate(&b));
  MLPTRY(m, hipEventRecord(a, m->stream));
  // State 5 where positive, the network on the other points of each 2048-point group.
  k_seed_block<<<(int)ceil_div(n, 2048), 256, 0, m->stream>>>(n, inc, incStride, state,
                                                               stateStride, m->theta, m->norm,
                                                               seed, true);
  MLPTRY(m, hipGetLastError());
  MLPTRY(m, hipEventRecord(b, m->stream));
  MLPTRY(m, hipEventSynchronize(b));
  float ms = 0.f;
  MLPTRY(m, hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (elapsedMs)
    *elapsedMs = ms;
  return VPD_OK;
}

} // extern "C"
